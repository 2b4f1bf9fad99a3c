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
      Stop 'dbs_gpu: library error'
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
