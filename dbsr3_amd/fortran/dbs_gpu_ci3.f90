!======================================================================
! dbs_gpu_ci3.f90 -- DBSR_CI3's accumulation path on the GPU with the
! reference's names and argument lists:
!     Call Add_integral(kpol,ic,jc,i1,i2,i3,i4,C,mbreit)
!                                          (DBSR_CI3/add_integral.f90:2-43)
!     Call Add_matrix(jtype,jpol)          (c_data, DBSR_MAT3/MOD_c_data.f90:318-368)
!     Call Gen_matrix(itype,kpol)          (DBSR_CI3/gen_matrix.f90:2-12)
!     Call Rk_evaluate(k,itype)            (DBSR_CI3/rk_evaluate.f90:2-63)
!     Call Sk_evaluate(k,itype)            (DBSR_CI3/sk_evaluate.f90:2-63)
! A host program written like DBSR_CI3's Setup_matrix
! (DBSR_CI3/setup_matrix.f90:133-156) links these instead of its own.
! Module dbs_gpu_ci plays the part of the module state those routines
! read (dbsr_ci: ibi, HM, DM, NZERO, eps_c; orb_jj: kef, lef;
! DBS_orbitals_pq: pq; c_data: the current merged list): the host hands it
! over once with dbs_gpu_ci_begin.  HM and DM stay the host's arrays; each
! Gen_matrix call brings them to the device, adds its list in list order
! and copies them back, so additions made by the host between flushes
! (Setup_matrix's met = 1 terms) keep the reference's order.
! Built into libdbsr3_amd_ci3.a and linked into the host program (DBSR_MAT3
! uses other routines under the same names: dbs_gpu_mat3.f90,
! libdbsr3_amd_mat3.a).
!======================================================================
      Module dbs_gpu_ci
      Use iso_c_binding
      Implicit none
      Type(c_ptr) :: cdata = c_null_ptr          ! the coefficient lists
      Type(c_ptr) :: pq_dev = c_null_ptr         ! pq(ns,2,norb) on the device
      Type(c_ptr) :: hm_dev = c_null_ptr, dm_dev = c_null_ptr
      Type(c_ptr) :: hm_host = c_null_ptr, dm_host = c_null_ptr
      Integer :: ibi = 2**15, norb = 0, ncj = 0, nzero = 0
      Real(8) :: eps_c = 1.d-10
      Integer(c_int), allocatable, target :: kef(:)   ! kappa by orbital, kef(0) unused
! ... the merged list handed to Gen_matrix (valid during the call)
      Integer :: nc_cur = 0
      Type(c_ptr) :: c_cur, k1_cur, k2_cur, k3_cur, k4_cur
      End Module dbs_gpu_ci

!----------------------------------------------------------------------
      Subroutine dbs_gpu_ci_gen(user, itype, kpol, nc, c, K1, K2, K3, K4) bind(C)
!----------------------------------------------------------------------
!     the library's Add_matrix hook: the merged list of (itype, kpol) in
!     IPT order, handed to the reference-named dispatcher Gen_matrix
      Use dbs_gpu_ci
      Implicit none
      Type(c_ptr), value :: user, c, K1, K2, K3, K4
      Integer(c_int), value :: itype, kpol, nc
      nc_cur = nc; c_cur = c; k1_cur = K1; k2_cur = K2; k3_cur = K3; k4_cur = K4
      Call Gen_matrix(itype, kpol)
      nc_cur = 0
      End Subroutine dbs_gpu_ci_gen

!----------------------------------------------------------------------
      Subroutine dbs_gpu_ci_begin(norb_, pq, kef_, ibi_, HM, ncj_, nzero_, DM, eps_c_, mtype, mpol, &
                                  mblock, nblock, kblock, eps_cdata)
!----------------------------------------------------------------------
!     hand the module state over: orbitals pq(ns,2,norb) (DBS_orbitals_pq),
!     kef(1:norb) (orb_jj), ibi, HM(ncj,nzero) and DM(ncj) (dbsr_ci, kept as
!     the host's arrays, which must stay in place until dbs_gpu_ci_end),
!     eps_c; then allocate the coefficient lists for
!     types 1..mtype and multipoles 0..mpol as alloc_c_data does
      Use dbs_gpu_api
      Use dbs_gpu_mem
      Use dbs_gpu_state, only: ctx, ns_g
      Use dbs_gpu_ci
      Implicit none
      Integer, intent(in) :: norb_, ibi_, ncj_, nzero_, mtype, mpol, mblock, nblock, kblock
      Real(8), intent(in), target :: pq(ns_g, 2, norb_)
      Integer, intent(in) :: kef_(norb_)
      Real(8), intent(inout), target :: HM(ncj_, nzero_), DM(ncj_)
      Real(8), intent(in) :: eps_c_, eps_cdata
      Interface
       Subroutine dbs_gpu_ci_gen(user, itype, kpol, nc, c, K1, K2, K3, K4) bind(C)
        Import :: c_ptr, c_int
        Type(c_ptr), value :: user, c, K1, K2, K3, K4
        Integer(c_int), value :: itype, kpol, nc
       End Subroutine
      End Interface
      if (.not. c_associated(ctx)) Stop 'dbs_gpu: call dbs_gpu_init first'
      Call dbs_gpu_ci_end
      norb = norb_; ibi = ibi_; ncj = ncj_; nzero = nzero_; eps_c = eps_c_
      if (allocated(kef)) Deallocate(kef)
      Allocate(kef(0:norb)); kef(0) = 0; kef(1:norb) = kef_
      pq_dev = dbs_gpu_put(pq, ns_g * 2 * norb)
      Call dbs_gpu_check(dbs_malloc(ctx, hm_dev, int(8 * max(ncj * nzero, 1), c_size_t)))
      Call dbs_gpu_check(dbs_malloc(ctx, dm_dev, int(8 * max(ncj, 1), c_size_t)))
      hm_host = c_loc(HM); dm_host = c_loc(DM)
      Call dbs_gpu_check(dbs_cdata_create(cdata, mtype, 0, mpol, mblock, nblock, kblock, eps_cdata, &
                         c_funloc(dbs_gpu_ci_gen), c_null_ptr))
      End Subroutine dbs_gpu_ci_begin

      Subroutine dbs_gpu_ci_end
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_ci
      Implicit none
      if (c_associated(cdata)) Call dbs_gpu_check(dbs_cdata_destroy(cdata))
      if (c_associated(pq_dev)) Call dbs_gpu_check(dbs_free(ctx, pq_dev))
      if (c_associated(hm_dev)) Call dbs_gpu_check(dbs_free(ctx, hm_dev))
      if (c_associated(dm_dev)) Call dbs_gpu_check(dbs_free(ctx, dm_dev))
      cdata = c_null_ptr; pq_dev = c_null_ptr; hm_dev = c_null_ptr; dm_dev = c_null_ptr
      End Subroutine dbs_gpu_ci_end

!======================================================================
      Subroutine Add_integral(kpol, ic, jc, i1, i2, i3, i4, C, mbreit)
!======================================================================
!     the term C*R^k(i1,i2;i3,i4) of HM(ic,jc): the parity rule, then the
!     four R^k types and, with mbreit, the eight S^k terms of the Breit
!     operator, into the coefficient lists (DBSR_CI3/add_integral.f90:2-43)
      Use dbs_gpu_api
      Use dbs_gpu_ci
      Implicit none
      Integer, intent(in) :: kpol, ic, jc, i1, i2, i3, i4, mbreit
      Real(8), intent(in) :: C
      Call dbs_gpu_check(dbs_add_integral_ci(cdata, kpol, ic, jc, i1, i2, i3, i4, C, kef, ibi, mbreit))
      End Subroutine Add_integral

!======================================================================
      Subroutine Add_matrix(jtype, jpol)
!======================================================================
!     merge the lists of (jtype, jpol) and hand them to Gen_matrix
!     (DBSR_MAT3/MOD_c_data.f90:318-368)
      Use dbs_gpu_api
      Use dbs_gpu_ci
      Implicit none
      Integer, intent(in) :: jtype, jpol
      Call dbs_gpu_check(dbs_cdata_add_matrix(cdata, jtype, jpol))
      End Subroutine Add_matrix

!======================================================================
      Subroutine Gen_matrix(itype, kpol)
!======================================================================
!     DBSR_CI3/gen_matrix.f90:2-12
      Implicit none
      Integer, intent(in) :: itype, kpol
      if (itype <= 4) then
       Call Rk_evaluate(kpol, itype)
      elseif (itype <= 6) then
       Call Sk_evaluate(kpol, itype)
      end if
      End Subroutine Gen_matrix

!======================================================================
      Subroutine Rk_evaluate(k, itype)
!======================================================================
!     every entry of the current merged list: S = its R^k (type itype) on
!     the GPU, HM(ic,jc) += S*c (jc <= NZERO) or DM(jc) += S*c (ic = jc),
!     entries with |c| < eps_c skipped, in list order per element
!     (DBSR_CI3/rk_evaluate.f90:2-63)
      Implicit none
      Integer, intent(in) :: k, itype
      if (itype < 1 .or. itype > 4) Stop 'Rk_evaluate: unknown type of intgrals'
      Call dbs_gpu_ci_evaluate(k, itype)
      End Subroutine Rk_evaluate

!======================================================================
      Subroutine Sk_evaluate(k, itype)
!======================================================================
!     the same for the S^k types 5 (ppqq) and 6 (pqqp)
!     (DBSR_CI3/sk_evaluate.f90:2-63)
      Implicit none
      Integer, intent(in) :: k, itype
      if (itype < 5 .or. itype > 6) Stop 'Sk_evaluate: unknown type of intgrals'
      Call dbs_gpu_ci_evaluate(k, itype)
      End Subroutine Sk_evaluate

      Subroutine dbs_gpu_ci_evaluate(k, itype)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_ci
      Implicit none
      Integer, intent(in) :: k, itype
      Integer(c_size_t) :: bh, bd
      if (nc_cur <= 0) Return
      bh = 8_c_size_t * ncj * nzero; bd = 8_c_size_t * ncj
      Call dbs_gpu_check(dbs_memcpy_h2d(ctx, hm_dev, hm_host, bh))
      Call dbs_gpu_check(dbs_memcpy_h2d(ctx, dm_dev, dm_host, bd))
      Call dbs_gpu_check(dbs_rk_evaluate_dev(ctx, itype, k, nc_cur, c_cur, k1_cur, k2_cur, k3_cur, k4_cur, &
                         eps_c, ibi, pq_dev, norb, hm_dev, ncj, nzero, dm_dev))
      Call dbs_gpu_check(dbs_memcpy_d2h(ctx, hm_host, hm_dev, bh))
      Call dbs_gpu_check(dbs_memcpy_d2h(ctx, dm_host, dm_dev, bd))
      End Subroutine dbs_gpu_ci_evaluate
