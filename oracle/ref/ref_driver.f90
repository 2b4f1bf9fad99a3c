!======================================================================
! ref_driver.f90 -- C-callable wrappers around the reference's own DBS
! library (DBSR_HF/dbsr_lib_dbs.f90 + dbsr_lib_zcom.f90) and its c_data
! module (DBSR_MAT3/MOD_c_data.f90), compiled from where those sources lie
! by oracle/Makefile into oracle/_ref/libdbsref.so.
!
! TEST INFRASTRUCTURE ONLY: used to pin the C restatement (dbs_oracle.c)
! against the reference's compiled code.  Never linked into the product.
!
! The wrappers set the reference's module state directly (DBS_grid) instead
! of going through def_grid, which reads knot files.
!======================================================================
      Subroutine ref_setup(ns_, ks_, ksp_, ksq_, t_) bind(C, name='ref_setup')
      Use iso_c_binding
      Use DBS_grid
      Implicit none
      Integer(c_int), value :: ns_, ks_, ksp_, ksq_
      Real(c_double), intent(in) :: t_(ns_ + ks_)
      ns = ns_; ks = ks_; ksp = ksp_; ksq = ksq_
      ms = ns + ns; nv = ns - ks + 1; nsp = nv + ksp - 1; nsq = nv + ksq - 1
      me = 0
      if (allocated(t)) Deallocate(t)
      Allocate(t(ns + ks)); t = t_
      tmax = t(ns + 1)
      Call alloc_DBS_gauss
      Call dealloc_DBS_integrals
      Call dealloc_DBS_sk_integrals
      Call alloc_DBS_moments
      End Subroutine ref_setup

      Subroutine ref_tables(gr_, grw_, pbsp_, qbsp_) bind(C, name='ref_tables')
      Use iso_c_binding
      Use DBS_grid
      Use DBS_gauss
      Implicit none
      Real(c_double) :: gr_(nv, ks), grw_(nv, ks), pbsp_(nv + 1, ks, ksp), qbsp_(nv + 1, ks, ksq)
      gr_ = gr; grw_ = grw; pbsp_ = pbsp; qbsp_ = qbsp
      End Subroutine ref_tables

      Subroutine ref_gauleg(n, x, w) bind(C, name='ref_gauleg')
      Use iso_c_binding
      Implicit none
      Integer(c_int), value :: n
      Real(c_double) :: x(n), w(n)
      Call gauleg(0.d0, 1.d0, x, w, n)
      End Subroutine ref_gauleg

      Subroutine ref_moments(which, k, rkm) bind(C, name='ref_moments')
      Use iso_c_binding
      Use DBS_grid
      Implicit none
      Integer(c_int), value :: which, k
      Real(c_double) :: rkm(ks * ks, nv)
      Select case (which)
       Case (0); Call moments_pp(k, ks * ks, nv, rkm)
       Case (1); Call moments_qq(k, ks * ks, nv, rkm)
       Case (2); Call moments_pq(k, ks * ks, nv, rkm)
       Case default; Call moments_qp(k, ks * ks, nv, rkm)
      End Select
      End Subroutine ref_moments

      Subroutine ref_triang(itype, k, iv, rkd_out) bind(C, name='ref_triang')
      Use iso_c_binding
      Use DBS_grid
      Use DBS_moments
      Implicit none
      Integer(c_int), value :: itype, k, iv
      Real(c_double) :: rkd_out(ks * ks, ks * ks)
      rkd(:, :, iv) = 0.d0
      Select case (itype)
       Case (1); Call triang_pppp(k, iv)
       Case (2); Call triang_qqqq(k, iv)
       Case (3); Call triang_pqpq(k, iv)
       Case (4); Call triang_qpqp(k, iv)
       Case (5); Call triang_ppqq(k, iv)
       Case default; Call triang_pqqp(k, iv)
      End Select
      rkd_out = rkd(:, :, iv)
      End Subroutine ref_triang

      Subroutine ref_select(itype, k) bind(C, name='ref_select')
      Use iso_c_binding
      Implicit none
      Integer(c_int), value :: itype, k
      Integer :: it, kk
      it = itype; kk = k
      Call ref_sel(it, kk)
      End Subroutine ref_select

      Subroutine ref_sel(itype, k)
      Implicit none
      Integer :: itype, k
      Select case (itype)
       Case (1); Call mrk_pppp(k)
       Case (2); Call mrk_qqqq(k)
       Case (3); Call mrk_pqpq(k)
       Case (4); Call mrk_qpqp(k)
       Case (5); Call msk_ppqq(k)
       Case default; Call msk_pqqp(k)
      End Select
      End Subroutine ref_sel

      Subroutine ref_mrk(itype, k, rkb_out) bind(C, name='ref_mrk')
      Use iso_c_binding
      Use DBS_grid
      Use DBS_integrals, only: rkb_r => rkb
      Use DBS_sk_integrals, only: rkb_s => rkb
      Implicit none
      Integer(c_int), value :: itype, k
      Real(c_double) :: rkb_out(*)
      Integer :: n
      Call ref_sel(int(itype), int(k))
      if (itype <= 4) then
       n = ns * ns * ks * ks
       rkb_out(1:n) = reshape(rkb_r, (/n/))
      else
       n = ns * ns * (2 * ks - 1) * (2 * ks - 1)
       rkb_out(1:n) = reshape(rkb_s, (/n/))
      end if
      End Subroutine ref_mrk

      Subroutine ref_convol(itype, k, icase, si, sj, d, a) bind(C, name='ref_convol')
      Use iso_c_binding
      Use DBS_grid
      Implicit none
      Integer(c_int), value :: itype, k, icase
      Character(kind=c_char), value :: si, sj
      Real(c_double) :: d(ns, *), a(ns, *)
      Character :: ci, cj
      ci = si; cj = sj
      Call ref_sel(int(itype), int(k))
      if (itype <= 4) then
       Call convol(ns, ks, a, d, icase, ci, cj)
      else
       Call convol_sk(ns, ks, a, d, icase, ci, cj)
      end if
      End Subroutine ref_convol

      Subroutine ref_density(p1, p2, typ, d) bind(C, name='ref_density')
      Use iso_c_binding
      Use DBS_grid
      Implicit none
      Real(c_double) :: p1(ns), p2(ns), d(ns, *)
      Character(kind=c_char), value :: typ
      Character :: c
      c = typ
      Call density(ns, ks, d, p1, p2, c)
      End Subroutine ref_density

      Function ref_sum_amb(a, b, sym) result(s) bind(C, name='ref_sum_amb')
      Use iso_c_binding
      Use DBS_grid
      Implicit none
      Real(c_double) :: a(ns, *), b(ns, *)
      Character(kind=c_char), value :: sym
      Real(c_double) :: s
      Character(1) :: c
      Real(8), external :: SUM_AmB
      c = sym
      s = SUM_AmB(ns, ks, a, b, c)
      End Function ref_sum_amb

      Function ref_rk(f1, f2, f3, f4, k) result(s) bind(C, name='ref_rk')
      Use iso_c_binding
      Use DBS_grid
      Implicit none
      Real(c_double) :: f1(ns, 2), f2(ns, 2), f3(ns, 2), f4(ns, 2)
      Integer(c_int), value :: k
      Real(c_double) :: s
      Real(8), external :: rk
      s = rk(f1, f2, f3, f4, k)
      End Function ref_rk

      Function ref_quadr(fi, fj, m) result(s) bind(C, name='ref_quadr')
      Use iso_c_binding
      Use DBS_grid
      Implicit none
      Real(c_double) :: fi(ns, 2), fj(ns, 2)
      Integer(c_int), value :: m
      Real(c_double) :: s
      Real(8), external :: quadr
      s = quadr(fi, fj, m)
      End Function ref_quadr

      Subroutine ref_update_hs(hm, ii, jj, d, sym) bind(C, name='ref_update_hs')
      Use iso_c_binding
      Use DBS_grid
      Implicit none
      Real(c_double) :: hm(ms, ms), d(ns, *)
      Integer(c_int), value :: ii, jj
      Character(kind=c_char), value :: sym
      Character(1) :: c
      c = sym
      Call UPDATE_HS(ms, hm, ii, jj, ns, ks, d, c)
      End Subroutine ref_update_hs

      Function ref_smu(ka, kb, kc, kd, L, v, S) result(r) bind(C, name='ref_smu')
      Use iso_c_binding
      Implicit none
      Integer(c_int), value :: ka, kb, kc, kd, L, v
      Real(c_double) :: S(8), r
      Real(8), external :: SMU
      Integer :: a, b, c, d, ll, vv
      a = ka; b = kb; c = kc; d = kd; ll = L; vv = v
      r = SMU(a, b, c, d, ll, vv, S)
      End Function ref_smu

      Subroutine ref_overlaps(fpbs_, fqbs_, fppqq_) bind(C, name='ref_overlaps')
      Use iso_c_binding
      Use DBS_grid
      Use DBS_gauss
      Implicit none
      Real(c_double) :: fpbs_(ns, ns), fqbs_(ns, ns), fppqq_(ms, ms)
      fpbs_ = fpbs; fqbs_ = fqbs; fppqq_ = fppqq
      End Subroutine ref_overlaps

      Subroutine ref_full_mat_sym(ns_, ks_, x, y, sym) bind(C, name='ref_full_mat_sym')
      Use iso_c_binding
      Implicit none
      Integer(c_int), value :: ns_, ks_
      Real(c_double) :: x(ns_, *), y(ns_, ns_)
      Character(kind=c_char), value :: sym
      Character(1) :: c
      c = sym
      Call Full_mat_sym(ns_, ks_, x, y, c)
      End Subroutine ref_full_mat_sym

      Subroutine ref_set_orbitals(n, pq) bind(C, name='ref_set_orbitals')
      Use iso_c_binding
      Use DBS_grid
      Use df_orbitals, only: p
      Implicit none
      Integer(c_int), value :: n
      Real(c_double), intent(in) :: pq(ns, 2, n)
      if (allocated(p)) Deallocate(p)
      Allocate(p(ns, 2, n)); p = pq
      End Subroutine ref_set_orbitals

      Function ref_sk(itype, i1, j1, i2, j2, k) result(s) bind(C, name='ref_sk')
      Use iso_c_binding
      Implicit none
      Integer(c_int), value :: itype, i1, j1, i2, j2, k
      Real(c_double) :: s
      Real(8), external :: sk_ppqq, sk_pqqp
      if (itype == 5) then
        s = sk_ppqq(i1, j1, i2, j2, k)
      else
        s = sk_pqqp(i1, j1, i2, j2, k)
      end if
      End Function ref_sk

      Subroutine ref_hf_setup(n, kmax_, mbreit_, kbs_, lbs_, jbs_, qsum_, coef_, pq, nrk_, k1, k2, k3, k4, c) &
                 bind(C, name='ref_hf_setup')
! ... module state of DBSR_HF for hf_matrix: orbitals, energy coefficients
! ... coef(i,j,k), kmax, mbreit and the Breit list (rk4_data)
      Use iso_c_binding
      Use DBS_grid
      Use DBS_dhl_pq
      Use df_orbitals
      Use hf_energy
      Use dbsr_hf, only: kmax, mbreit
      Use rk4_data
      Implicit none
      Integer(c_int), value :: n, kmax_, mbreit_, nrk_
      Integer(c_int), intent(in) :: kbs_(n), lbs_(n), jbs_(n), k1(*), k2(*), k3(*), k4(*)
      Real(c_double), intent(in) :: qsum_(n), coef_(n, n, 0:kmax_), pq(ns, 2, n), c(*)
      Call alloc_df_orbitals(n)
      Call alloc_DF_radial(ns)
      kbs(1:n) = kbs_; lbs(1:n) = lbs_; jbs(1:n) = jbs_; qsum(1:n) = qsum_
      p(:, :, 1:n) = pq
      if (allocated(coef)) Deallocate(coef)
      Allocate(coef(n, n, 0:kmax_)); coef = coef_
      kmax = kmax_; mbreit = mbreit_
      if (allocated(crk)) Deallocate(crk, kr1, kr2, kr3, kr4)
      mrk = max(nrk_, 1); Allocate(crk(mrk), kr1(mrk), kr2(mrk), kr3(mrk), kr4(mrk))
      nrk = nrk_
      if (nrk > 0) then
        crk(1:nrk) = c(1:nrk); kr1(1:nrk) = k1(1:nrk); kr2(1:nrk) = k2(1:nrk)
        kr3(1:nrk) = k3(1:nrk); kr4(1:nrk) = k4(1:nrk)
      end if
! ... the one-electron operator is outside the two-electron path: dhl = 0
      Call alloc_dhl_pq(ms)
      End Subroutine ref_hf_setup

      Subroutine ref_hf_matrix(i, hfm) bind(C, name='ref_hf_matrix')
      Use iso_c_binding
      Use DBS_grid
      Use DBS_dhl_pq
      Use df_orbitals, only: kbs
      Implicit none
      Integer(c_int), value :: i
      Real(c_double) :: hfm(ms, ms)
      k_dhl = kbs(i)
      Call hf_matrix(i, hfm)
      End Subroutine ref_hf_matrix

      Function ref_mkgrid(gtype, ks_, ml_, hi_, he_, hmax_, rmax_, z, nucl, afermi, runiform, ns_in, t_, tcap) &
               result(nsout) bind(C, name='ref_mkgrid')
! ... the reference's mkgrid_01 (gtype 1) / mkgrid_02 (gtype 2) with the
! ... nucleus of DBS_nuclear: nucl 0 'point', 1 'Fermi' (a_fermi), 2
! ... 'uniform' (r_uniform); mkgrid_02 takes ns_in.  Returns ns and, when
! ... tcap >= ns+ks, the knots.
      Use iso_c_binding
      Use DBS_grid
      Use DBS_nuclear
      Implicit none
      Integer(c_int), value :: gtype, ks_, ml_, nucl, ns_in, tcap
      Real(c_double), value :: hi_, he_, hmax_, rmax_, z, afermi, runiform
      Real(c_double) :: t_(*)
      Integer(c_int) :: nsout
      ks = ks_; ml = ml_; hi = hi_; he = he_; hmax = hmax_; rmax = rmax_
      atomic_number = z; a_fermi = afermi; r_uniform = runiform
      nuclear = 'point'
      if (nucl == 1) nuclear = 'Fermi'
      if (nucl == 2) nuclear = 'uniform'
      if (allocated(t)) Deallocate(t)
      if (gtype == 1) then
        Call mkgrid_01
      else
        ns = ns_in
        Call mkgrid_02
      end if
      nsout = ns
      if (tcap >= ns + ks) t_(1:ns + ks) = t(1:ns + ks)
      End Function ref_mkgrid

      Function ref_read_grid(path, plen, ns_, ks_, t_) result(rc) bind(C, name='ref_read_grid')
! ... the reference's read_grid on a knot file (grid points listed after
! ... 'grid points:'); ns_ and ks_ as the file states them
      Use iso_c_binding
      Use DBS_grid
      Implicit none
      Integer(c_int), value :: plen, ns_, ks_
      Character(kind=c_char) :: path(plen)
      Real(c_double) :: t_(ns_ + ks_)
      Integer(c_int) :: rc
      Character(512) :: fname
      Integer :: i, nu
      fname = ' '
      Do i = 1, plen
       fname(i:i) = path(i)
      End do
      nu = 77
      Open(nu, file=trim(fname), status='old')
      ns = ns_; ks = ks_
      Call read_grid(nu)
      Close(nu)
      t_ = t(1:ns + ks)
      rc = 0
      End Function ref_read_grid

      Function ref_breit_energy(eps) result(e) bind(C, name='ref_breit_energy')
! ... the reference's Breit_energy (DBSR_HF/hf_energy_BR.f90:2-50) for the
! ... state set by ref_hf_setup (orbitals, kbs, coef(i,j,k), kmax); eps is
! ... its cut on the merged coefficients (module dbsr_hf eps_c)
      Use iso_c_binding
      Use dbsr_hf, only: eps_c
      Use rk4_data
      Implicit none
      Real(c_double), value :: eps
      Real(c_double) :: e
      Real(8) :: eb
      eps_c = eps
      nrk = 0
      Call Breit_energy(eb)
      e = eb
      End Function ref_breit_energy

      Subroutine ref_rk_rows(n, rows, norb, pq, out) bind(C, name='ref_rk_rows')
! ... R^k of the rows (k,a,b,c,d) (5,n), sorted by k and left pair (a,c),
! ... with the reference's own routines in Rk_evaluate's organisation
! ... (DBSR_CI3/rk_evaluate.f90:31-56) for the rk() type sum
! ... (dbsr_lib_dbs.f90:3981-4022): per type and multipole mrk_*, one
! ... density + convol case 2 per left pair, density + SUM_AmB per row;
! ... the four type values of a row added in rk()'s order
      Use iso_c_binding
      Use DBS_grid, only: ns, ks
      Implicit none
      Integer(c_int), value :: n, norb
      Integer(c_int), intent(in) :: rows(5, n)
      Real(c_double), intent(in) :: pq(ns, 2, norb)
      Real(c_double) :: out(n)
      Real(8), allocatable :: part(:,:), d1(:,:), d2(:,:), cv(:,:)
      Real(8), external :: SUM_AmB
      Integer :: it, i, k, kcur, la, lc, cl, cr
      Integer, parameter :: compl(4) = (/1, 2, 2, 1/), compr(4) = (/1, 2, 1, 2/)
      Allocate(part(4, n), d1(ns, ks), d2(ns, ks), cv(ns, ks))
      Do it = 1, 4
       kcur = -1; la = -1; lc = -1
       cl = compl(it); cr = compr(it)
       Do i = 1, n
        k = rows(1, i)
        if (k /= kcur) then
         Select case(it)
          Case(1); Call mrk_pppp(k)
          Case(2); Call mrk_qqqq(k)
          Case(3); Call mrk_qpqp(k)
          Case(4); Call mrk_pqpq(k)
         End select
         kcur = k; la = -1
        end if
        if (rows(2, i) /= la .or. rows(4, i) /= lc) then
         la = rows(2, i); lc = rows(4, i)
         Call density(ns, ks, d1, pq(1, cl, la), pq(1, cl, lc), 's')
         Call convol(ns, ks, cv, d1, 2, 's', 's')
        end if
        Call density(ns, ks, d2, pq(1, cr, rows(3, i)), pq(1, cr, rows(5, i)), 's')
        part(it, i) = SUM_AmB(ns, ks, cv, d2, 's')
       End do
      End do
      Do i = 1, n
       out(i) = 0.d0 + part(1, i)
       out(i) = out(i) + part(2, i)
       out(i) = out(i) + part(3, i)
       out(i) = out(i) + part(4, i)
      End do
      Deallocate(part, d1, d2, cv)
      End Subroutine ref_rk_rows

      Subroutine ref_set_nucleus(model, z, b, a, c) bind(C, name='ref_set_nucleus')
! ... module DBS_nuclear: 0 point, 1 uniform (radius b), 2 Fermi (a, c)
      Use iso_c_binding
      Use DBS_nuclear
      Implicit none
      Integer(c_int), value :: model
      Real(c_double), value :: z, b, a, c
      atomic_number = z; r_uniform = b; a_fermi = a; c_fermi = c
      if (model == 0) nuclear = 'point'
      if (model == 1) nuclear = 'uniform'
      if (model == 2) nuclear = 'Fermi'
      End Subroutine ref_set_nucleus

      Function ref_vnuclear(r) result(v) bind(C, name='ref_vnuclear')
      Use iso_c_binding
      Implicit none
      Real(c_double), value :: r
      Real(c_double) :: v
      Real(8), external :: V_nuclear
      v = V_nuclear(r)
      End Function ref_vnuclear

      Subroutine ref_dhl_pq(kappa, out) bind(C, name='ref_dhl_pq')
! ... Def_Vnucl, then MAT_dhl_pq(kappa) with the module's Bloch defaults
      Use iso_c_binding
      Use DBS_grid
      Use DBS_dhl_pq
      Implicit none
      Integer(c_int), value :: kappa
      Real(c_double) :: out(ms, ms)
      Call Def_Vnucl
      Call alloc_dhl_pq(ms)
      Call MAT_dhl_pq(kappa)
      out = dhl
      End Subroutine ref_dhl_pq

      Subroutine ref_av_energy(nconf_, iqconf_, weight_, out) bind(C, name='ref_av_energy')
! ... DBSR_HF's average-term energy of the orbitals set by ref_hf_setup:
! ... configurations iqconf(nconf,nbf) with weights, qsum from them
! ... (Read_confs), av_energy_coef, update_int(0) (Vp_dhl, hf_fk, hf_gk),
! ... ENERGY; then the configuration energies as Conf_energies forms them.
! ... out = (Etotal, E1body, E2body, E(conf 1..nconf))
      Use iso_c_binding
      Use DBS_grid
      Use DBS_dhl_pq
      Use df_orbitals
      Use hf_energy
      Use dbsr_hf, only: nconf, iqconf, weight, kmax, Etotal, E1body, E2body
      Implicit none
      Integer(c_int), value :: nconf_
      Integer(c_int), intent(in) :: iqconf_(nconf_, nbf)
      Real(c_double), intent(in) :: weight_(nconf_)
      Real(c_double) :: out(3 + nconf_)
      Integer :: io, ic
      nconf = nconf_
      if (allocated(iqconf)) Deallocate(iqconf)
      if (allocated(weight)) Deallocate(weight)
      Allocate(iqconf(nconf, nbf), weight(nconf))
      iqconf = iqconf_; weight = weight_ / SUM(weight_)
      qsum = 0.d0
      Do io = 1, nbf
       Do ic = 1, nconf
        qsum(io) = qsum(io) + iqconf(ic, io) * weight(ic)
       End do
      End do
      kmax = maxval(jbs(1:nbf))
      if (allocated(coef)) Deallocate(coef)
      if (allocated(value)) Deallocate(value)
      if (allocated(valueI)) Deallocate(valueI)
      Allocate(coef(nbf, nbf, 0:kmax), value(nbf, nbf, 0:kmax), valueI(nbf))
      coef = 0.d0; value = 0.d0; valueI = 0.d0
      Call Def_Vnucl
      Call alloc_dhl_pq(ms)
      Call av_energy_coef(nbf)
      Call update_int(0)
      Call energy
      out(1) = Etotal; out(2) = E1body; out(3) = E2body
      Do ic = 1, nconf
       qsum(:) = iqconf(ic, :)
       weight = 0.d0; weight(ic) = 1.d0; coef = 0.d0
       Call av_energy_coef(nbf)
       Call energy
       out(3 + ic) = Etotal
      End do
      End Subroutine ref_av_energy

      Subroutine ref_av_tables(kmx, vals, vali, coefo) bind(C, name='ref_av_tables')
! ... module hf_energy's tables after ref_av_energy: value(nbf,nbf,0:kmax),
! ... valueI(nbf) and coef(nbf,nbf,0:kmax) of the last configuration set
      Use iso_c_binding
      Use df_orbitals, only: nbf
      Use hf_energy
      Implicit none
      Integer(c_int), value :: kmx
      Real(c_double) :: vals(nbf, nbf, 0:kmx), vali(nbf), coefo(nbf, nbf, 0:kmx)
      vals = value; vali = valueI; coefo = coef
      End Subroutine ref_av_tables
