!======================================================================
! drv_dropin.f90 -- a host program written the way the reference's callers
! are (rk_evaluate.f90, hf_rk.f90): it calls mrk_*, density, Convol,
! SUM_AmB, rk and quadr by their reference names and gets the GPU versions
! from dbsr3_amd/fortran/dbs_gpu.f90; then hf_matrix, the engine, and
! mchf_matrix / core_matrix / hf_eiv with their module data passed
! explicitly.  Reads drv_in.bin, writes drv_out.bin (stream files
! written/read by tests/test_gpu_fortran.py).
!======================================================================
      Program drv_dropin
      Implicit none
      Integer :: ns, ks, ksp, ksq, norb, k, n, i
      Real(8), allocatable :: t(:), pq(:,:,:), d1(:,:), d2(:,:), conv(:,:), out(:)
      Real(8), allocatable :: dn1(:,:), dn2(:,:), convn(:,:)
      Real(8), allocatable :: acoef(:,:), bcoef(:,:), hfm(:,:), grad(:,:,:), crk(:)
      Integer, allocatable :: rows(:,:), kr(:)
      Real(8) :: energy, coef(3)
      Integer :: nl, nr
      Integer, allocatable :: l_i(:,:), r_i(:,:), iprm(:)
      Real(8), allocatable :: l_c(:), r_c(:), dhl(:,:), fppqq(:,:), mh(:,:), mr(:), mx(:,:), ch(:,:), cx(:,:), &
                              qs(:), vo(:,:), ev(:), he(:,:)
      Real(8) :: e_i, q_i, e_v
      Real(8), external :: rk, SUM_AmB, quadr

      Open(11, file='drv_in.bin', access='stream', form='unformatted', status='old')
      Read(11) ns, ks, ksp, ksq
      Allocate(t(ns + ks)); Read(11) t
      Read(11) norb, k
      Allocate(pq(ns, 2, norb)); Read(11) pq
      Read(11) nl, nr
      Allocate(l_i(2, nl), l_c(nl), r_i(5, nr), r_c(nr), dhl(2 * ns, 2 * ns), fppqq(2 * ns, 2 * ns), iprm(2 * ns))
      Read(11) l_i; Read(11) l_c; Read(11) r_i; Read(11) r_c; Read(11) dhl; Read(11) fppqq; Read(11) iprm
      Close(11)

      Call dbs_gpu_init(ns, ks, ksp, ksq, t)
      Allocate(d1(ns, ks), d2(ns, ks), conv(ns, ks), out(10))
      Allocate(dn1(ns, 2*ks-1), dn2(ns, 2*ks-1), convn(ns, 2*ks-1))

! ... R^k(1,2;3,4) as rk() assembles it
      out(1) = rk(pq(1,1,1), pq(1,1,2), pq(1,1,3), pq(1,1,4), k)

! ... the same integral the way Rk_evaluate does it, one type at a time
      out(2) = 0.d0
      Call mrk_pppp(k)
      Call Density(ns, ks, d1, pq(1,1,1), pq(1,1,3), 's')
      Call Convol(ns, ks, conv, d1, 2, 's', 's')
      Call Density(ns, ks, d2, pq(1,1,2), pq(1,1,4), 's')
      out(2) = out(2) + SUM_AmB(ns, ks, conv, d2, 's')
      Call mrk_qqqq(k)
      Call Density(ns, ks, d1, pq(1,2,1), pq(1,2,3), 's')
      Call Convol(ns, ks, conv, d1, 2, 's', 's')
      Call Density(ns, ks, d2, pq(1,2,2), pq(1,2,4), 's')
      out(2) = out(2) + SUM_AmB(ns, ks, conv, d2, 's')
      Call mrk_qpqp(k)
      Call Density(ns, ks, d1, pq(1,2,1), pq(1,2,3), 's')
      Call Convol(ns, ks, conv, d1, 2, 's', 's')
      Call Density(ns, ks, d2, pq(1,1,2), pq(1,1,4), 's')
      out(2) = out(2) + SUM_AmB(ns, ks, conv, d2, 's')
      Call mrk_pqpq(k)
      Call Density(ns, ks, d1, pq(1,1,1), pq(1,1,3), 's')
      Call Convol(ns, ks, conv, d1, 2, 's', 's')
      Call Density(ns, ks, d2, pq(1,2,2), pq(1,2,4), 's')
      out(2) = out(2) + SUM_AmB(ns, ks, conv, d2, 's')

! ... <1|r^m|2> for m = 0, 1, -1
      out(3) = quadr(pq(1,1,1), pq(1,1,2), 0)
      out(4) = quadr(pq(1,1,1), pq(1,1,2), 1)
      out(5) = quadr(pq(1,1,1), pq(1,1,2), -1)
! ... a direct potential (convol case 1) contracted back: <d1|V(d2)>
      Call mrk_pppp(k)
      Call Density(ns, ks, d1, pq(1,1,1), pq(1,1,3), 's')
      Call Density(ns, ks, d2, pq(1,1,2), pq(1,1,4), 's')
      Call Convol(ns, ks, conv, d2, 1, 's', 's')
      out(6) = SUM_AmB(ns, ks, conv, d1, 's')
      out(7) = dble(ns); out(8) = dble(k)
! ... interleaved selections, as sk_ppqq (DBSR_HF/hf_sk.f90) and a direct
!     potential would mix them: convol reads the last mrk_* array,
!     convol_sk the last msk_* array, whichever was selected last
      Call msk_ppqq(k)
      Call mrk_pppp(k)
      Call Density(ns, ks, dn1, pq(1,1,1), pq(1,2,3), 'n')
      Call Convol_sk(ns, ks, convn, dn1, 2, 'n', 'n')
      Call Density(ns, ks, dn2, pq(1,1,2), pq(1,2,4), 'n')
      out(9) = SUM_AmB(ns, ks, convn, dn2, 'n')
      Call mrk_qqqq(k)
      Call msk_pqqp(k)
      Call Density(ns, ks, d1, pq(1,2,1), pq(1,2,3), 's')
      Call Density(ns, ks, d2, pq(1,2,2), pq(1,2,4), 's')
      Call Convol(ns, ks, conv, d1, 2, 's', 's')
      out(10) = SUM_AmB(ns, ks, conv, d2, 's')

! ... hf_matrix's two-electron potential of orbital 1: a(1,2,0) = 1,
!     b(1,3,1) = 0.5, qsum = 2, no Breit terms
      Allocate(acoef(norb, 0:2), bcoef(norb, 0:2), hfm(2 * ns, 2 * ns), kr(1), crk(1))
      acoef = 0.d0; bcoef = 0.d0; acoef(2, 0) = 1.d0; bcoef(3, 1) = 0.5d0
      hfm = 0.d0; kr = 0; crk = 0.d0
      Call dbs_gpu_hf_matrix(1, norb, pq, acoef, bcoef, 2, 2.d0, 0, 0, kr, kr, kr, kr, crk, hfm)
! ... the engine: E = sum c R^k and dE/d(pq) for three rows, twice (a
!     captured graph replayed with new coefficients)
      Allocate(rows(5, 3), grad(ns, 2, norb))
      rows(:, 1) = (/0, 1, 2, 3, 4/); rows(:, 2) = (/1, 1, 1, 2, 2/); rows(:, 3) = (/2, 2, 3, 2, 3/)
      Call dbs_gpu_engine_setup(norb, 3, rows)
      coef = (/1.d0, -0.5d0, 0.25d0/)
      Call dbs_gpu_engine_step(pq, 2.d0 * coef, energy, grad)
      Call dbs_gpu_engine_step(pq, coef, energy, grad)
      Call dbs_gpu_engine_close

! ... mchf_matrix of orbital 1 (newton), hf_eiv of qsum*dhl + hf_matrix's
!     potential, orthogonal to orbital 2, and core_matrix of orbital 1
!     (newton) with acoef(1,1) = 0.75 added and qsum = 2, 1.5, 1, 1
      Allocate(mh(2 * ns, 2 * ns), mr(2 * ns), mx(2 * ns, 2 * ns), ch(2 * ns, 2 * ns), cx(2 * ns, 2 * ns), &
               qs(norb), vo(2 * ns, 1), ev(2 * ns), he(2 * ns, 2 * ns))
      Call dbs_gpu_mchf_matrix(1, norb, pq, dhl, nl, l_i(1, :), l_i(2, :), l_c, nr, r_i(1, :), r_i(2, :), &
                               r_i(3, :), r_i(4, :), r_i(5, :), r_c, 2, 1.d-8, 1.d-8, 1, 0, fppqq, mh, mr, mx, &
                               e_i, q_i)
      he = 2.d0 * dhl
      Call dbs_gpu_hf_matrix(1, norb, pq, acoef, bcoef, 2, 2.d0, 0, 0, kr, kr, kr, kr, crk, he)
      vo(1:ns, 1) = pq(:, 1, 2); vo(ns + 1:2 * ns, 1) = pq(:, 2, 2)
      Call dbs_gpu_hf_eiv(he, 1, vo, iprm, 1, fppqq, ev, e_v)
      qs = 1.d0; qs(1) = 2.d0; qs(2) = 1.5d0
      acoef(1, 1) = 0.75d0
      Call dbs_gpu_core_matrix(1, norb, pq, dhl, acoef, bcoef, 2, qs, 1, ch, cx)

      Open(12, file='drv_out.bin', access='stream', form='unformatted', status='replace')
      Write(12) out
      Write(12) hfm
      Write(12) energy
      Write(12) grad
      Write(12) mh, mr, mx, e_i, q_i
      Write(12) ch, cx
      Write(12) ev, e_v
      Close(12)
      Call dbs_gpu_close
      End Program drv_dropin
