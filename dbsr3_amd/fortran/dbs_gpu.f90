!======================================================================
! dbs_gpu.f90 -- Fortran side of the drop-in boundary.
!
! Module dbs_gpu_api declares the C-ABI of libdbsr3_amd.so (include/
! dbsr3_amd.h) with ISO_C_BINDING.  The external procedures below carry
! the names, argument lists and meaning of the reference's DBS library
! routines (DBSR_HF/dbsr_lib_dbs.f90), so that a host program calling
!     Call mrk_pppp(k); Call Convol(ns,ks,a,d,2,'s','s'); S = SUM_AmB(...)
! runs them on the GPU when it links these instead of the library's own
! definitions.  Module dbs_gpu_state plays the part of the reference's
! module state (DBS_grid sizes and the two rkb pointers): it holds the
! context, the (type,k) that the last mrk_* selected (DBS_integrals::rkb,
! read by convol) and the one the last msk_* selected
! (DBS_sk_integrals::rkb, read by convol_sk).
! A nonzero status from the library stops with the library's message, as
! the reference's Stop statements do.
!======================================================================
      Module dbs_gpu_api
      Use iso_c_binding
      Implicit none
      Interface
       Function dbs_ctx_create(ctx, ns, ks, ksp, ksq, t, device) bind(C, name='dbs_ctx_create')
        Import :: c_ptr, c_int, c_double
        Type(c_ptr) :: ctx
        Integer(c_int), value :: ns, ks, ksp, ksq, device
        Real(c_double) :: t(*)
        Integer(c_int) :: dbs_ctx_create
       End Function
       Function dbs_ctx_destroy(ctx) bind(C, name='dbs_ctx_destroy')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: ctx
        Integer(c_int) :: dbs_ctx_destroy
       End Function
       Function dbs_last_error() bind(C, name='dbs_last_error')
        Import :: c_ptr
        Type(c_ptr) :: dbs_last_error
       End Function
       Function dbs_mrk_host(ctx, itype, k, rkb) bind(C, name='dbs_mrk_host')
        Import :: c_ptr, c_int, c_double
        Type(c_ptr), value :: ctx
        Integer(c_int), value :: itype, k
        Real(c_double) :: rkb(*)
        Integer(c_int) :: dbs_mrk_host
       End Function
       Function dbs_mrk(ctx, itype, k, rkb_dev) bind(C, name='dbs_mrk')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: ctx
        Integer(c_int), value :: itype, k
        Type(c_ptr) :: rkb_dev
        Integer(c_int) :: dbs_mrk
       End Function
       Function dbs_density(ctx, p1, p2, typ, d) bind(C, name='dbs_density')
        Import :: c_ptr, c_int, c_double, c_char
        Type(c_ptr), value :: ctx
        Real(c_double) :: p1(*), p2(*), d(*)
        Character(kind=c_char), value :: typ
        Integer(c_int) :: dbs_density
       End Function
       Function dbs_convol(ctx, itype, k, icase, si, sj, d, a) bind(C, name='dbs_convol')
        Import :: c_ptr, c_int, c_double, c_char
        Type(c_ptr), value :: ctx
        Integer(c_int), value :: itype, k, icase
        Character(kind=c_char), value :: si, sj
        Real(c_double) :: d(*), a(*)
        Integer(c_int) :: dbs_convol
       End Function
       Function dbs_sum_amb(ctx, a, b, sym, out) bind(C, name='dbs_sum_amb')
        Import :: c_ptr, c_int, c_double, c_char
        Type(c_ptr), value :: ctx
        Real(c_double) :: a(*), b(*), out
        Character(kind=c_char), value :: sym
        Integer(c_int) :: dbs_sum_amb
       End Function
       Function dbs_rk(ctx, f1, f2, f3, f4, k, out) bind(C, name='dbs_rk')
        Import :: c_ptr, c_int, c_double
        Type(c_ptr), value :: ctx
        Real(c_double) :: f1(*), f2(*), f3(*), f4(*), out
        Integer(c_int), value :: k
        Integer(c_int) :: dbs_rk
       End Function
       Function dbs_quadr(ctx, fi, fj, m, out) bind(C, name='dbs_quadr')
        Import :: c_ptr, c_int, c_double
        Type(c_ptr), value :: ctx
        Real(c_double) :: fi(*), fj(*), out
        Integer(c_int), value :: m
        Integer(c_int) :: dbs_quadr
       End Function
       Function dbs_ctx_sync(ctx) bind(C, name='dbs_ctx_sync')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: ctx
        Integer(c_int) :: dbs_ctx_sync
       End Function
       Function dbs_malloc(ctx, p, bytes) bind(C, name='dbs_malloc')
        Import :: c_ptr, c_int, c_size_t
        Type(c_ptr), value :: ctx
        Type(c_ptr) :: p
        Integer(c_size_t), value :: bytes
        Integer(c_int) :: dbs_malloc
       End Function
       Function dbs_free(ctx, p) bind(C, name='dbs_free')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: ctx, p
        Integer(c_int) :: dbs_free
       End Function
       Function dbs_memcpy_h2d(ctx, dst, src, bytes) bind(C, name='dbs_memcpy_h2d')
        Import :: c_ptr, c_int, c_size_t
        Type(c_ptr), value :: ctx, dst, src
        Integer(c_size_t), value :: bytes
        Integer(c_int) :: dbs_memcpy_h2d
       End Function
       Function dbs_memcpy_d2h(ctx, dst, src, bytes) bind(C, name='dbs_memcpy_d2h')
        Import :: c_ptr, c_int, c_size_t
        Type(c_ptr), value :: ctx, dst, src
        Integer(c_size_t), value :: bytes
        Integer(c_int) :: dbs_memcpy_d2h
       End Function
! ... coefficient lists (module c_data) and Add_integral
       Function dbs_cdata_create(cd, ntype, kpol1, kpol2, mblock, nblock, kblock, eps_c, gen, user) &
                bind(C, name='dbs_cdata_create')
        Import :: c_ptr, c_int, c_double, c_funptr
        Type(c_ptr) :: cd
        Integer(c_int), value :: ntype, kpol1, kpol2, mblock, nblock, kblock
        Real(c_double), value :: eps_c
        Type(c_funptr), value :: gen
        Type(c_ptr), value :: user
        Integer(c_int) :: dbs_cdata_create
       End Function
       Function dbs_cdata_destroy(cd) bind(C, name='dbs_cdata_destroy')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: cd
        Integer(c_int) :: dbs_cdata_destroy
       End Function
       Function dbs_cdata_add_coef(cd, c, kpol, j1, j2, j3, j4, itype) bind(C, name='dbs_cdata_add_coef')
        Import :: c_ptr, c_int, c_double
        Type(c_ptr), value :: cd
        Real(c_double), value :: c
        Integer(c_int), value :: kpol, j1, j2, j3, j4, itype
        Integer(c_int) :: dbs_cdata_add_coef
       End Function
       Function dbs_cdata_add_matrix(cd, itype, kpol) bind(C, name='dbs_cdata_add_matrix')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: cd
        Integer(c_int), value :: itype, kpol
        Integer(c_int) :: dbs_cdata_add_matrix
       End Function
       Function dbs_add_integral_ci(cd, kpol, ic, jc, i1, i2, i3, i4, c, kappa, ibi, mbreit) &
                bind(C, name='dbs_add_integral_ci')
        Import :: c_ptr, c_int, c_double
        Type(c_ptr), value :: cd
        Integer(c_int), value :: kpol, ic, jc, i1, i2, i3, i4, ibi, mbreit
        Real(c_double), value :: c
        Integer(c_int) :: kappa(*)
        Integer(c_int) :: dbs_add_integral_ci
       End Function
       Function dbs_add_integral_mat3(cd, orbs, icase, kpol, i1, i2, i3, i4, c, ic, jc, io, jo) &
                bind(C, name='dbs_add_integral_mat3')
        Import :: c_ptr, c_int, c_double
        Type(c_ptr), value :: cd, orbs
        Integer(c_int), value :: icase, kpol, i1, i2, i3, i4, ic, jc, io, jo
        Real(c_double), value :: c
        Integer(c_int) :: dbs_add_integral_mat3
       End Function
! ... list evaluation and accumulation
       Function dbs_rk_evaluate_dev(ctx, itype, k, nc, c, K1, K2, K3, K4, eps_c, ibi, pq_dev, norb, &
                HM_dev, ldhm, nzero, DM_dev) bind(C, name='dbs_rk_evaluate_dev')
        Import :: c_ptr, c_int, c_double
        Type(c_ptr), value :: ctx, c, K1, K2, K3, K4, pq_dev, HM_dev, DM_dev
        Integer(c_int), value :: itype, k, nc, ibi, norb, ldhm, nzero
        Real(c_double), value :: eps_c
        Integer(c_int) :: dbs_rk_evaluate_dev
       End Function
       Function dbs_rdata_dev(ctx, m, aux, itype, atype, k, nc, c, K1, K2, K3, K4, pq_dev, norb) &
                bind(C, name='dbs_rdata_dev')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: ctx, m, aux, c, K1, K2, K3, K4, pq_dev
        Integer(c_int), value :: itype, atype, k, nc, norb
        Integer(c_int) :: dbs_rdata_dev
       End Function
       Function dbs_sdata_dev(ctx, m, aux, ktype, k, nc, c, K1, K2, K3, K4, pq_dev, norb) &
                bind(C, name='dbs_sdata_dev')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: ctx, m, aux, c, K1, K2, K3, K4, pq_dev
        Integer(c_int), value :: ktype, k, nc, norb
        Integer(c_int) :: dbs_sdata_dev
       End Function
       Function dbs_o_data_dev(ctx, m, itype, n, k1, k2, c, eps_c, ibo, ipbs, kbort, kbort_user, vch_dev, &
                nvch, fppqq_dev, iitar, iptar, otarg) bind(C, name='dbs_o_data_dev')
        Import :: c_ptr, c_int, c_double, c_funptr
        Type(c_ptr), value :: ctx, m, k1, k2, c, ipbs, kbort_user, vch_dev, fppqq_dev, iptar, otarg
        Type(c_funptr), value :: kbort
        Integer(c_int), value :: itype, n, ibo, nvch, iitar
        Real(c_double), value :: eps_c
        Integer(c_int) :: dbs_o_data_dev
       End Function
       Function dbs_ldata_dev(ctx, m, aux, itype, n, k1, k2, k3, k4, c, eps_c, pq_dev) &
                bind(C, name='dbs_ldata_dev')
         Import :: c_int, c_ptr, c_double
         Type(c_ptr), value :: ctx, m, aux, k1, k2, k3, k4, c, pq_dev
         Integer(c_int), value :: itype, n
         Real(c_double), value :: eps_c
         Integer(c_int) :: dbs_ldata_dev
       End Function
       Function dbs_mchf_matrix_dev(ctx, i, norb, pq_dev, dhl_dev, nl, l_i1, l_i2, l_c, nr, r_k, r_i1, r_i2, &
                r_i3, r_i4, r_c, kmax, eps_c, qsum_min, newton, irhs, fppqq_dev, hfm_dev, rhs_dev, hx_dev, &
                e_i, qsum_i) bind(C, name='dbs_mchf_matrix_dev')
         Import :: c_int, c_ptr, c_double
         Type(c_ptr), value :: ctx, pq_dev, dhl_dev, l_i1, l_i2, l_c, r_k, r_i1, r_i2, r_i3, r_i4, r_c, &
                               fppqq_dev, hfm_dev, rhs_dev, hx_dev
         Integer(c_int), value :: i, norb, nl, nr, kmax, newton, irhs
         Real(c_double), value :: eps_c, qsum_min
         Real(c_double) :: e_i, qsum_i
         Integer(c_int) :: dbs_mchf_matrix_dev
       End Function
       Function dbs_core_matrix_dev(ctx, i, norb, pq_dev, hd_dev, acoef, bcoef, kmax, qsum, newton, hfm_dev, &
                hx_dev) bind(C, name='dbs_core_matrix_dev')
         Import :: c_int, c_ptr
         Type(c_ptr), value :: ctx, pq_dev, hd_dev, acoef, bcoef, qsum, hfm_dev, hx_dev
         Integer(c_int), value :: i, norb, kmax, newton
         Integer(c_int) :: dbs_core_matrix_dev
       End Function
       Function dbs_hf_eiv_dev(ctx, hfm_dev, northo, vorth_dev, iprm, m, fppqq_dev, v_dev, e_i) &
                bind(C, name='dbs_hf_eiv_dev')
         Import :: c_int, c_ptr, c_double
         Type(c_ptr), value :: ctx, hfm_dev, vorth_dev, iprm, fppqq_dev, v_dev
         Integer(c_int), value :: northo, m
         Real(c_double) :: e_i
         Integer(c_int) :: dbs_hf_eiv_dev
       End Function
       Function dbs_hf_matrix_dev(ctx, i, norb, pq_dev, acoef, bcoef, kmax, qsum, mbreit, nrk, kr1, kr2, &
                kr3, kr4, crk, hfm_dev) bind(C, name='dbs_hf_matrix_dev')
        Import :: c_ptr, c_int, c_double
        Type(c_ptr), value :: ctx, pq_dev, acoef, bcoef, kr1, kr2, kr3, kr4, crk, hfm_dev
        Integer(c_int), value :: i, norb, kmax, mbreit, nrk
        Real(c_double), value :: qsum
        Integer(c_int) :: dbs_hf_matrix_dev
       End Function
! ... channel-block post-processing (DBSR_MAT3)
       Function dbs_zero_cond(nch, ns, nsp, nsq, ksp, ksq, lch, kch, ilzero, jlzero, ibzero, jbzero, npert, &
                iprm, nhm) bind(C, name='dbs_zero_cond')
        Import :: c_ptr, c_int
        Integer(c_int), value :: nch, ns, nsp, nsq, ksp, ksq, ilzero, jlzero, ibzero, jbzero, npert
        Type(c_ptr), value :: lch, kch, iprm
        Integer(c_int) :: nhm
        Integer(c_int) :: dbs_zero_cond
       End Function
       Function dbs_orth_conditions(nch, ipch, nbf, ipbs, kbs, ibort, ibort_user, ldo, nort, iort) &
                bind(C, name='dbs_orth_conditions')
        Import :: c_ptr, c_int, c_funptr
        Integer(c_int), value :: nch, nbf, ldo
        Type(c_ptr), value :: ipch, ipbs, kbs, ibort_user, nort, iort
        Type(c_funptr), value :: ibort
        Integer(c_int) :: dbs_orth_conditions
       End Function
       Function dbs_orth_dev(ctx, m, idiag, nort, iort, ldo, pv_dev, qv_dev, nbf) bind(C, name='dbs_orth_dev')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: ctx, m, nort, iort, pv_dev, qv_dev
        Integer(c_int), value :: idiag, ldo, nbf
        Integer(c_int) :: dbs_orth_dev
       End Function
       Function dbs_diag_channels_dev(ctx, m, diag_dev, nd, iprm, edmin, edmax, egap, ipsol) &
                bind(C, name='dbs_diag_channels_dev')
        Import :: c_ptr, c_int, c_double
        Type(c_ptr), value :: ctx, m, diag_dev, iprm, ipsol
        Integer(c_int), value :: nd
        Real(c_double), value :: edmin, edmax, egap
        Integer(c_int) :: dbs_diag_channels_dev
       End Function
       Function dbs_record_matrix_dev(ctx, m, diag_dev, nd, ipsol, c_dev, c_cap, coff, w_dev, w_cap, woff, &
                overlaps) bind(C, name='dbs_record_matrix_dev')
        Import :: c_ptr, c_int, c_long_long
        Type(c_ptr), value :: ctx, m, diag_dev, ipsol, c_dev, coff, w_dev, woff, overlaps
        Integer(c_int), value :: nd
        Integer(c_long_long), value :: c_cap, w_cap
        Integer(c_int) :: dbs_record_matrix_dev
       End Function
! ... the batched engine
       Function dbs_engine_create(eng, ninst, ctxs, norb, nrows, rows, kind) bind(C, name='dbs_engine_create')
        Import :: c_ptr, c_int
        Type(c_ptr) :: eng
        Integer(c_int), value :: ninst, kind
        Type(c_ptr) :: ctxs(*)
        Integer(c_int) :: norb(*), nrows(*), rows(*)
        Integer(c_int) :: dbs_engine_create
       End Function
       Function dbs_engine_destroy(eng) bind(C, name='dbs_engine_destroy')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: eng
        Integer(c_int) :: dbs_engine_destroy
       End Function
       Function dbs_engine_buffers(eng, inst, pq, coef, vals, grad, energy) bind(C, name='dbs_engine_buffers')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: eng
        Integer(c_int), value :: inst
        Type(c_ptr) :: pq, coef, vals, grad, energy
        Integer(c_int) :: dbs_engine_buffers
       End Function
       Function dbs_engine_forward_adjoint(eng) bind(C, name='dbs_engine_forward_adjoint')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: eng
        Integer(c_int) :: dbs_engine_forward_adjoint
       End Function
       Function dbs_engine_graph_capture(eng, with_adjoint) bind(C, name='dbs_engine_graph_capture')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: eng
        Integer(c_int), value :: with_adjoint
        Integer(c_int) :: dbs_engine_graph_capture
       End Function
       Function dbs_engine_graph_launch(eng) bind(C, name='dbs_engine_graph_launch')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: eng
        Integer(c_int) :: dbs_engine_graph_launch
       End Function
       Function dbs_engine_sync(eng) bind(C, name='dbs_engine_sync')
        Import :: c_ptr, c_int
        Type(c_ptr), value :: eng
        Integer(c_int) :: dbs_engine_sync
       End Function
      End Interface
      End Module dbs_gpu_api

      Module dbs_gpu_state
      Use iso_c_binding
      Implicit none
      Type(c_ptr) :: ctx = c_null_ptr
      Integer :: ns_g = 0, ks_g = 0
      Integer :: itype_r = 0, krk_r = -100   ! last mrk_* (R^k array)
      Integer :: itype_s = 0, krk_s = -100   ! last msk_* (S^k array)
      End Module dbs_gpu_state

!----------------------------------------------------------------------
      Subroutine dbs_gpu_check(rc)
!----------------------------------------------------------------------
!     stop with the library's message on a nonzero status
      Use dbs_gpu_api
      Implicit none
      Integer(c_int), intent(in) :: rc
      Character(kind=c_char), pointer :: msg(:)
      Character(256) :: text
      Integer :: i
      if (rc == 0) Return
      Call c_f_pointer(dbs_last_error(), msg, (/256/))
      text = ' '
      Do i = 1, 256
       if (msg(i) == c_null_char) Exit
       text(i:i) = msg(i)
      End do
      write(*, '(a)') trim(text)
      Error stop 'dbs_gpu: library error'
      End Subroutine dbs_gpu_check

!----------------------------------------------------------------------
      Subroutine dbs_gpu_init(ns, ks, ksp, ksq, t)
!----------------------------------------------------------------------
!     create the context from the grid of module DBS_grid (call after
!     def_grid / alloc_DBS_gauss with DBS_grid's ns, ks, ksp, ksq, t)
      Use dbs_gpu_api
      Use dbs_gpu_state
      Implicit none
      Integer, intent(in) :: ns, ks, ksp, ksq
      Real(8), intent(in) :: t(ns + ks)
      Real(c_double) :: tt(ns + ks)
      if (c_associated(ctx)) Call dbs_gpu_check(dbs_ctx_destroy(ctx))
      tt = t
      Call dbs_gpu_check(dbs_ctx_create(ctx, ns, ks, ksp, ksq, tt, 0))
      ns_g = ns; ks_g = ks
      itype_r = 0; krk_r = -100; itype_s = 0; krk_s = -100
      End Subroutine dbs_gpu_init

      Subroutine dbs_gpu_close
      Use dbs_gpu_api
      Use dbs_gpu_state
      Implicit none
      if (c_associated(ctx)) Call dbs_gpu_check(dbs_ctx_destroy(ctx))
      ctx = c_null_ptr
      End Subroutine dbs_gpu_close

!----------------------------------------------------------------------
!     mrk_pppp(k) ... msk_pqqp(k): select (and build on the device) the
!     array of (type,k), as the reference points rkb at rka(...,k,type)
!----------------------------------------------------------------------
      Subroutine dbs_gpu_select(itype, k)
      Use dbs_gpu_api
      Use dbs_gpu_state
      Implicit none
      Integer, intent(in) :: itype, k
      Type(c_ptr) :: p
      if (.not. c_associated(ctx)) Stop 'dbs_gpu: call dbs_gpu_init first'
      Call dbs_gpu_check(dbs_mrk(ctx, itype, k, p))
      if (itype <= 4) then
       itype_r = itype; krk_r = k
      else
       itype_s = itype; krk_s = k
      end if
      End Subroutine dbs_gpu_select

      Subroutine mrk_pppp(k)
      Integer, intent(in) :: k
      Call dbs_gpu_select(1, k)
      End Subroutine mrk_pppp

      Subroutine mrk_qqqq(k)
      Integer, intent(in) :: k
      Call dbs_gpu_select(2, k)
      End Subroutine mrk_qqqq

      Subroutine mrk_pqpq(k)
      Integer, intent(in) :: k
      Call dbs_gpu_select(3, k)
      End Subroutine mrk_pqpq

      Subroutine mrk_qpqp(k)
      Integer, intent(in) :: k
      Call dbs_gpu_select(4, k)
      End Subroutine mrk_qpqp

      Subroutine msk_ppqq(k)
      Integer, intent(in) :: k
      Call dbs_gpu_select(5, k)
      End Subroutine msk_ppqq

      Subroutine msk_pqqp(k)
      Integer, intent(in) :: k
      Call dbs_gpu_select(6, k)
      End Subroutine msk_pqqp

!----------------------------------------------------------------------
      Subroutine density(ns, ks, d, p1, p2, type)
!----------------------------------------------------------------------
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Implicit none
      Integer :: ns, ks
      Real(8) :: d(ns, *), p1(ns), p2(ns)
      Character :: type
      Call dbs_gpu_check(dbs_density(ctx, p1, p2, type, d))
      End Subroutine density

!----------------------------------------------------------------------
      Subroutine convol(ns, ks, a, d, icase, sym_i, sym_j)
!----------------------------------------------------------------------
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx, itype_r, krk_r
      Implicit none
      Integer, intent(in) :: ns, ks, icase
      Character, intent(in) :: sym_i, sym_j
      Real(8), intent(in) :: d(ns, *)
      Real(8), intent(out) :: a(ns, *)
      if (itype_r == 0) Stop 'convol: no array selected'
      if (icase < 1 .or. icase > 4) Stop 'convol: unknown case'
      Call dbs_gpu_check(dbs_convol(ctx, itype_r, krk_r, icase, sym_i, sym_j, d, a))
      End Subroutine convol

!----------------------------------------------------------------------
      Subroutine convol_sk(ns, ks, a, d, icase, sym_i, sym_j)
!----------------------------------------------------------------------
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx, itype_s, krk_s
      Implicit none
      Integer, intent(in) :: ns, ks, icase
      Character, intent(in) :: sym_i, sym_j
      Real(8), intent(in) :: d(ns, *)
      Real(8), intent(out) :: a(ns, *)
      if (itype_s == 0) Stop 'convol_sk: no array selected'
      if (icase < 1 .or. icase > 4) Stop 'convol: unknown case'
      Call dbs_gpu_check(dbs_convol(ctx, itype_s, krk_s, icase, sym_i, sym_j, d, a))
      End Subroutine convol_sk

!----------------------------------------------------------------------
      Real(8) Function SUM_AmB(ns, ks, a, b, sym)
!----------------------------------------------------------------------
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Implicit none
      Integer, intent(in) :: ns, ks
      Real(8), intent(in) :: a(ns, *), b(ns, *)
      Character(*), intent(in) :: sym
      Real(c_double) :: x
      Call dbs_gpu_check(dbs_sum_amb(ctx, a, b, sym(1:1), x))
      SUM_AmB = x
      End Function SUM_AmB

!----------------------------------------------------------------------
      Real(8) Function rk(f1, f2, f3, f4, k)
!----------------------------------------------------------------------
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx, ns_g
      Implicit none
      Integer, intent(in) :: k
      Real(8), intent(in) :: f1(ns_g, 2), f2(ns_g, 2), f3(ns_g, 2), f4(ns_g, 2)
      Real(c_double) :: x
      Call dbs_gpu_check(dbs_rk(ctx, f1, f2, f3, f4, k, x))
      rk = x
      End Function rk

!----------------------------------------------------------------------
      Real(8) Function quadr(fi, fj, m)
!----------------------------------------------------------------------
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx, ns_g
      Implicit none
      Integer, intent(in) :: m
      Real(8), intent(in) :: fi(ns_g, 2), fj(ns_g, 2)
      Real(c_double) :: x
      Call dbs_gpu_check(dbs_quadr(ctx, fi, fj, m, x))
      quadr = x
      End Function quadr

!----------------------------------------------------------------------
!     device buffers for the routines below (context stream)
!----------------------------------------------------------------------
      Module dbs_gpu_mem
      Use iso_c_binding
      Implicit none
      Contains

      Function dbs_gpu_put(a, n) result(p)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Integer, intent(in) :: n
      Real(8), intent(in), target :: a(n)
      Type(c_ptr) :: p
      Call dbs_gpu_check(dbs_malloc(ctx, p, int(8 * max(n, 1), c_size_t)))
      if (n > 0) Call dbs_gpu_check(dbs_memcpy_h2d(ctx, p, c_loc(a), 8_c_size_t * n))
      End Function dbs_gpu_put

      Subroutine dbs_gpu_get(p, a, n)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Type(c_ptr), intent(in) :: p
      Integer, intent(in) :: n
      Real(8), intent(out), target :: a(n)
      if (n > 0) Call dbs_gpu_check(dbs_memcpy_d2h(ctx, c_loc(a), p, 8_c_size_t * n))
      End Subroutine dbs_gpu_get

      End Module dbs_gpu_mem

!----------------------------------------------------------------------
      Subroutine dbs_gpu_hf_matrix(i, norb, pq, acoef, bcoef, kmax, qsum, mbreit, nrk, kr1, kr2, kr3, kr4, &
                                   crk, hfm)
!----------------------------------------------------------------------
!     the two-electron part of hf_matrix(i,hfm) (DBSR_HF/hf_matrix.f90:2-78,
!     Breit block :97-163 when mbreit = 2) on the GPU, with the module data
!     the reference reads passed explicitly: orbitals pq(ns,2,norb), the
!     coefficients a(i,j,k), b(i,j,k) as acoef/bcoef(norb,0:kmax), qsum(i)
!     and the Breit list of rk4_data.  hfm(ms,ms) holds on entry what the
!     host put there (qsum(i)*dhl in the reference) and on return the
!     potential, divided by qsum, as hf_matrix leaves it.
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx, ns_g
      Use dbs_gpu_mem
      Implicit none
      Integer, intent(in) :: i, norb, kmax, mbreit, nrk
      Real(8), intent(in), target :: pq(ns_g, 2, norb), acoef(norb, 0:kmax), bcoef(norb, 0:kmax), qsum
      Integer, intent(in), target :: kr1(*), kr2(*), kr3(*), kr4(*)
      Real(8), intent(in), target :: crk(*)
      Real(8), intent(inout), target :: hfm(2 * ns_g, 2 * ns_g)
      Type(c_ptr) :: pqd, hd
      pqd = dbs_gpu_put(pq, ns_g * 2 * norb)
      hd = dbs_gpu_put(hfm, 4 * ns_g * ns_g)
      Call dbs_gpu_check(dbs_hf_matrix_dev(ctx, i, norb, pqd, c_loc(acoef), c_loc(bcoef), kmax, qsum, mbreit, &
                         nrk, c_loc(kr1), c_loc(kr2), c_loc(kr3), c_loc(kr4), c_loc(crk), hd))
      Call dbs_gpu_get(hd, hfm, 4 * ns_g * ns_g)
      Call dbs_gpu_check(dbs_free(ctx, pqd))
      Call dbs_gpu_check(dbs_free(ctx, hd))
      End Subroutine dbs_gpu_hf_matrix

!----------------------------------------------------------------------
      Subroutine dbs_gpu_mchf_matrix(i, norb, pq, dhl, nl, l_i1, l_i2, l_c, nr, r_k, r_i1, r_i2, r_i3, r_i4, &
                                     r_c, kmax, eps_c, qsum_min, newton, irhs, fppqq, hfm, rhs, hx, e_i, qsum_i)
!----------------------------------------------------------------------
!     mchf_matrix(i,hfm,rhs,hx) (DBSR_MCHF/mchf_matrix.f90:2-248) on the GPU
!     with the module data passed explicitly: orbitals pq(ns,2,norb),
!     dhl = Get_hd_core(kbs(i)), the L and R^k lists flattened per angular
!     coefficient with c = WW(ic,jc)*coef, eps_c, qsum_min, newton, irhs,
!     fppqq.  Returns hfm, rhs, hx, e(i) and qsum(i).
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx, ns_g
      Use dbs_gpu_mem
      Implicit none
      Integer, intent(in) :: i, norb, nl, nr, kmax, newton, irhs
      Integer(c_int), intent(in), target :: l_i1(*), l_i2(*), r_k(*), r_i1(*), r_i2(*), r_i3(*), r_i4(*)
      Real(8), intent(in), target :: pq(ns_g, 2, norb), dhl(2 * ns_g, 2 * ns_g), l_c(*), r_c(*), &
                                     fppqq(2 * ns_g, 2 * ns_g)
      Real(8), intent(in) :: eps_c, qsum_min
      Real(8), intent(out), target :: hfm(2 * ns_g, 2 * ns_g), rhs(2 * ns_g), hx(2 * ns_g, 2 * ns_g)
      Real(8), intent(out) :: e_i, qsum_i
      Type(c_ptr) :: pqd, dd, bd, hd, rd, xd
      Integer :: mm
      mm = 4 * ns_g * ns_g
      pqd = dbs_gpu_put(pq, ns_g * 2 * norb); dd = dbs_gpu_put(dhl, mm); bd = dbs_gpu_put(fppqq, mm)
      hd = dbs_gpu_put(hfm, mm); rd = dbs_gpu_put(rhs, 2 * ns_g); xd = dbs_gpu_put(hx, mm)
      Call dbs_gpu_check(dbs_mchf_matrix_dev(ctx, i, norb, pqd, dd, nl, c_loc(l_i1), c_loc(l_i2), c_loc(l_c), &
                         nr, c_loc(r_k), c_loc(r_i1), c_loc(r_i2), c_loc(r_i3), c_loc(r_i4), c_loc(r_c), kmax, &
                         eps_c, qsum_min, newton, irhs, bd, hd, rd, xd, e_i, qsum_i))
      Call dbs_gpu_get(hd, hfm, mm); Call dbs_gpu_get(rd, rhs, 2 * ns_g); Call dbs_gpu_get(xd, hx, mm)
      Call dbs_gpu_check(dbs_free(ctx, pqd)); Call dbs_gpu_check(dbs_free(ctx, dd))
      Call dbs_gpu_check(dbs_free(ctx, bd)); Call dbs_gpu_check(dbs_free(ctx, hd))
      Call dbs_gpu_check(dbs_free(ctx, rd)); Call dbs_gpu_check(dbs_free(ctx, xd))
      End Subroutine dbs_gpu_mchf_matrix

!----------------------------------------------------------------------
      Subroutine dbs_gpu_core_matrix(i, norb, pq, hd, acoef, bcoef, kmax, qsum, newton, hfm, hx)
!----------------------------------------------------------------------
!     core_matrix(i,hfm,hx) (DBSR_MCHF/mchf_matrix.f90:253-345) on the GPU:
!     hd = Get_hd(kbs(i)), a(i,j,k) / b(i,j,k) as acoef/bcoef(norb,0:kmax),
!     qsum(norb)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx, ns_g
      Use dbs_gpu_mem
      Implicit none
      Integer, intent(in) :: i, norb, kmax, newton
      Real(8), intent(in), target :: pq(ns_g, 2, norb), hd(2 * ns_g, 2 * ns_g), acoef(norb, 0:kmax), &
                                     bcoef(norb, 0:kmax), qsum(norb)
      Real(8), intent(out), target :: hfm(2 * ns_g, 2 * ns_g), hx(2 * ns_g, 2 * ns_g)
      Type(c_ptr) :: pqd, dd, hfd, xd
      Integer :: mm
      mm = 4 * ns_g * ns_g
      pqd = dbs_gpu_put(pq, ns_g * 2 * norb); dd = dbs_gpu_put(hd, mm)
      hfd = dbs_gpu_put(hfm, mm); xd = dbs_gpu_put(hx, mm)
      Call dbs_gpu_check(dbs_core_matrix_dev(ctx, i, norb, pqd, dd, c_loc(acoef), c_loc(bcoef), kmax, &
                         c_loc(qsum), newton, hfd, xd))
      Call dbs_gpu_get(hfd, hfm, mm); Call dbs_gpu_get(xd, hx, mm)
      Call dbs_gpu_check(dbs_free(ctx, pqd)); Call dbs_gpu_check(dbs_free(ctx, dd))
      Call dbs_gpu_check(dbs_free(ctx, hfd)); Call dbs_gpu_check(dbs_free(ctx, xd))
      End Subroutine dbs_gpu_core_matrix

!----------------------------------------------------------------------
      Subroutine dbs_gpu_hf_eiv(hfm, northo, vorth, iprm, m, fppqq, v, e_i)
!----------------------------------------------------------------------
!     hf_eiv(i,hfm,v) (DBSR_HF/hf_solve_HF.f90:118-202) on the GPU with the
!     module data passed explicitly: the orbitals to project out
!     vorth(ms,northo) (the lower orbitals of the same kappa), iprm(ms) of
!     Boundary_conditions, m = n - l - northo.  Returns v(ms) and e(i,i).
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx, ns_g
      Use dbs_gpu_mem
      Implicit none
      Integer, intent(in) :: northo, m
      Integer(c_int), intent(in), target :: iprm(2 * ns_g)
      Real(8), intent(in), target :: hfm(2 * ns_g, 2 * ns_g), vorth(2 * ns_g, *), fppqq(2 * ns_g, 2 * ns_g)
      Real(8), intent(out), target :: v(2 * ns_g)
      Real(8), intent(out) :: e_i
      Type(c_ptr) :: hd, od, bd, vd
      Integer :: mm
      mm = 4 * ns_g * ns_g
      hd = dbs_gpu_put(hfm, mm); bd = dbs_gpu_put(fppqq, mm); vd = dbs_gpu_put(v, 2 * ns_g)
      od = dbs_gpu_put(vorth(:, 1:max(northo, 1)), 2 * ns_g * max(northo, 1))
      Call dbs_gpu_check(dbs_hf_eiv_dev(ctx, hd, northo, od, c_loc(iprm), m, bd, vd, e_i))
      Call dbs_gpu_get(vd, v, 2 * ns_g)
      Call dbs_gpu_check(dbs_free(ctx, hd)); Call dbs_gpu_check(dbs_free(ctx, bd))
      Call dbs_gpu_check(dbs_free(ctx, vd)); Call dbs_gpu_check(dbs_free(ctx, od))
      End Subroutine dbs_gpu_hf_eiv

!----------------------------------------------------------------------
!     the batched engine for an SCF loop: E = sum_n coef_n R^k(rows_n)
!     and dE/d(pq), the direct+exchange potentials applied to the
!     orbitals (what hf_matrix builds), replayed as one captured hipGraph
!     per call.  rows(5,nrows) = (k,a,b,c,d), 1-based orbitals.
!----------------------------------------------------------------------
      Module dbs_gpu_engine_state
      Use iso_c_binding
      Implicit none
      Type(c_ptr) :: eng = c_null_ptr, pq_d = c_null_ptr, coef_d = c_null_ptr, grad_d = c_null_ptr, &
                     e_d = c_null_ptr
      Integer :: norb_e = 0, nrows_e = 0
      End Module dbs_gpu_engine_state

      Subroutine dbs_gpu_engine_setup(norb, nrows, rows)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_engine_state
      Implicit none
      Integer, intent(in) :: norb, nrows
      Integer(c_int), intent(in) :: rows(5, nrows)
      Type(c_ptr) :: ctxs(1), vals
      Integer(c_int) :: no(1), nr(1)
      if (c_associated(eng)) Call dbs_gpu_check(dbs_engine_destroy(eng))
      ctxs(1) = ctx; no(1) = norb; nr(1) = nrows
      Call dbs_gpu_check(dbs_engine_create(eng, 1, ctxs, no, nr, rows, 0))
      Call dbs_gpu_check(dbs_engine_buffers(eng, 0, pq_d, coef_d, vals, grad_d, e_d))
      Call dbs_gpu_check(dbs_engine_graph_capture(eng, 1))
      norb_e = norb; nrows_e = nrows
      End Subroutine dbs_gpu_engine_setup

      Subroutine dbs_gpu_engine_step(pq, coef, energy, grad)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx, ns_g
      Use dbs_gpu_engine_state
      Implicit none
      Real(8), intent(in), target :: pq(ns_g, 2, norb_e), coef(nrows_e)
      Real(8), intent(out), target :: energy, grad(ns_g, 2, norb_e)
      if (.not. c_associated(eng)) Stop 'dbs_gpu_engine_step: call dbs_gpu_engine_setup first'
      Call dbs_gpu_check(dbs_memcpy_h2d(ctx, pq_d, c_loc(pq), int(8 * ns_g * 2 * norb_e, c_size_t)))
      Call dbs_gpu_check(dbs_memcpy_h2d(ctx, coef_d, c_loc(coef), int(8 * nrows_e, c_size_t)))
      Call dbs_gpu_check(dbs_ctx_sync(ctx))
      Call dbs_gpu_check(dbs_engine_graph_launch(eng))
      Call dbs_gpu_check(dbs_engine_sync(eng))
      Call dbs_gpu_check(dbs_memcpy_d2h(ctx, c_loc(energy), e_d, int(8, c_size_t)))
      Call dbs_gpu_check(dbs_memcpy_d2h(ctx, c_loc(grad), grad_d, int(8 * ns_g * 2 * norb_e, c_size_t)))
      End Subroutine dbs_gpu_engine_step

      Subroutine dbs_gpu_engine_close
      Use dbs_gpu_api
      Use dbs_gpu_engine_state
      Implicit none
      if (c_associated(eng)) Call dbs_gpu_check(dbs_engine_destroy(eng))
      eng = c_null_ptr
      End Subroutine dbs_gpu_engine_close
