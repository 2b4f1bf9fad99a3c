!======================================================================
! dbs_gpu_mat3_post.f90 -- DBSR_MAT3's channel post-processing on the GPU
! with the reference's names (a member of libdbsr3_amd_mat3.a of its own,
! so that host programs without it need not define its externals):
!     Call Zero_cond                  (DBSR_MAT3/zero_cond.f90:2-56)
!     Call DBS_ORTH                   (DBSR_MAT3/dbs_orth.f90:2-84)
!     Call Diag_channels              (DBSR_MAT3/diag.f90:2-79)
!     Call Record_matrix              (DBSR_MAT3/record_matrix.f90:2-60)
! on the storage of module dbs_gpu_mat (dbs_gpu_mat3.f90), with the data
! handed over by dbs_gpu_mat_channels and dbs_gpu_mat_save_diag; the host
! program's IBORT(i,j), Get_pv(j,v,ns) and Get_qv(j,v,ns) are called for
! the orthogonality conditions, as the reference does.
!======================================================================
      Function dbs_gpu_ibort(user, i, j) bind(C) result(k)
      Use iso_c_binding
      Implicit none
      Type(c_ptr), value :: user
      Integer(c_int), value :: i, j
      Integer(c_int) :: k
      Integer, external :: IBORT
      k = IBORT(i, j)
      End Function dbs_gpu_ibort

!======================================================================
      Subroutine Zero_cond
!======================================================================
!     DBSR_MAT3/zero_cond.f90:2-56: iprm(nch*ms+npert), nhm
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ns_g
      Use dbs_gpu_mat
      Implicit none
      if (.not. allocated(lch)) Stop 'Zero_cond: call dbs_gpu_mat_channels first'
      if (allocated(iprm)) Deallocate(iprm)
      Allocate(iprm(nch_c * ms + mb%npert))
      Call dbs_gpu_check(dbs_zero_cond(nch_c, ns_g, nsp_c, nsq_c, ksp_c, ksq_c, c_loc(lch), c_loc(kch), ilzero, &
                         jlzero, ibzero, jbzero, mb%npert, c_loc(iprm), nhm))
      End Subroutine Zero_cond

!======================================================================
      Subroutine DBS_ORTH
!======================================================================
!     DBSR_MAT3/dbs_orth.f90:2-84: the conditions by the reference's rule
!     with the host's IBORT, the orbital vectors from the host's Get_pv /
!     Get_qv, the projections of the blocks idiag selects on the device
      Use dbs_gpu_api
      Use dbs_gpu_mem
      Use dbs_gpu_state, only: ctx, ns_g
      Use dbs_gpu_mat
      Implicit none
      Integer(c_int), allocatable, target :: nort(:), iort(:)
      Real(8), allocatable, target :: pv(:,:), qv(:,:)
      Logical, allocatable :: used(:)
      Type(c_ptr) :: pvd, qvd
      Integer :: j, ich
      Interface
       Function dbs_gpu_ibort(user, i, j) bind(C) result(k)
        Import :: c_ptr, c_int
        Type(c_ptr), value :: user
        Integer(c_int), value :: i, j
        Integer(c_int) :: k
       End Function
      End Interface
      if (.not. allocated(lch)) Stop 'DBS_ORTH: call dbs_gpu_mat_channels first'
      Allocate(nort(max(nch_c, 1)), iort(max(nch_c, 1) * max(norb, 1)), used(max(norb, 1)))
      Call dbs_gpu_check(dbs_orth_conditions(nch_c, c_loc(ipch), norb, c_loc(ipbs), c_loc(kbs), &
                         c_funloc(dbs_gpu_ibort), c_null_ptr, norb, c_loc(nort), c_loc(iort)))
      if (nch_c == 0 .or. maxval(nort(1:max(nch_c, 1))) == 0) Return
      used = .false.
      Do ich = 1, nch_c
       Do j = 1, nort(ich); used(iort((ich - 1) * norb + j)) = .true.; End do
      End do
      Allocate(pv(ms, norb), qv(ms, norb)); pv = 0.d0; qv = 0.d0
      Do j = 1, norb
       if (.not. used(j)) Cycle
       Call Get_pv(j, pv(:, j), ns_g)
       Call Get_qv(j, qv(:, j), ns_g)
      End do
      pvd = dbs_gpu_put(pv, ms * norb); qvd = dbs_gpu_put(qv, ms * norb)
      Call dbs_gpu_check(dbs_orth_dev(ctx, c_loc(mb), idiag, c_loc(nort), c_loc(iort), norb, pvd, qvd, norb))
      Call dbs_gpu_check(dbs_free(ctx, pvd)); Call dbs_gpu_check(dbs_free(ctx, qvd))
      End Subroutine DBS_ORTH

!======================================================================
      Subroutine Diag_channels
!======================================================================
!     DBSR_MAT3/diag.f90:2-79: Zero_cond, then the diagonal blocks'
!     eigen-problems against the overlaps kept by dbs_gpu_mat_save_diag;
!     the solutions replace them, ipsol(0:nch)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_mat
      Implicit none
      if (.not. c_associated(diag_dev)) Stop 'Diag_channels: call dbs_gpu_mat_save_diag first'
      Call Zero_cond
      if (allocated(ipsol)) Deallocate(ipsol)
      Allocate(ipsol(0:nch_c))
      Call dbs_gpu_check(dbs_diag_channels_dev(ctx, c_loc(mb), diag_dev, nch_c, c_loc(iprm), Edmin, Edmax, Egap, &
                         c_loc(ipsol)))
      End Subroutine Diag_channels

!======================================================================
      Subroutine Record_matrix
!======================================================================
!     DBSR_MAT3/record_matrix.f90:2-60: the blocks of the current storage
!     transformed to the channel solutions on the device, written to unit
!     nui in the reference's records; overlaps for icase 0
      Use dbs_gpu_api
      Use dbs_gpu_mem
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_mat
      Implicit none
      Integer(c_long_long), allocatable, target :: coff(:), woff(:)
      Real(8), allocatable, target :: c(:), w(:), hpv(:)
      Type(c_ptr) :: cd, wd, ovp
      Integer(c_long_long) :: nc, nw, o
      Integer :: ich, jch, ip, jp, k, ni, nj, np
      if (.not. allocated(ipsol)) Stop 'Record_matrix: call Diag_channels first'
      np = mb%npert
      nc = 0; nw = 0
      Do ich = 2, nch_c; Do jch = 1, ich - 1
       if (icc(ich, jch) /= 0) nc = nc + int(ipsol(ich), c_long_long) * ipsol(jch)
      End do; End do
      Do ich = 1, nch_c; Do ip = 1, np
       if (icb(ich, ip) /= 0) nw = nw + ipsol(ich)
      End do; End do
      Allocate(coff(max(nch_c * nch_c, 1)), woff(max(nch_c * np, 1)), c(max(nc, 1_c_long_long)), &
               w(max(nw, 1_c_long_long)), hpv(max(mb%nbb, 1)))
      Call dbs_gpu_check(dbs_malloc(ctx, cd, int(8, c_size_t) * max(nc, 1_c_long_long)))
      Call dbs_gpu_check(dbs_malloc(ctx, wd, int(8, c_size_t) * max(nw, 1_c_long_long)))
      ovp = c_null_ptr
      if (icase == 0) then
       if (allocated(overlaps)) Deallocate(overlaps)
       Allocate(overlaps(nch_c + np, nch_c + np)); overlaps = 0.d0
       ovp = c_loc(overlaps)
      end if
      Call dbs_gpu_check(dbs_record_matrix_dev(ctx, c_loc(mb), diag_dev, nch_c, c_loc(ipsol), cd, nc, c_loc(coff), &
                         wd, nw, c_loc(woff), ovp))
      Call dbs_gpu_get(cd, c, int(nc)); Call dbs_gpu_get(wd, w, int(nw))
      Call dbs_gpu_get(mb%hp_dev, hpv, mb%nbb)
      Call dbs_gpu_check(dbs_free(ctx, cd)); Call dbs_gpu_check(dbs_free(ctx, wd))
      Do ich = 2, nch_c; ni = ipsol(ich)
      Do jch = 1, ich - 1; nj = ipsol(jch)
       o = coff(ich + nch_c * (jch - 1)); if (o < 0) Cycle
       write(nui) ich, jch
       write(nui) c(o + 1:o + int(ni, c_long_long) * nj)
      End do; End do
      if (np > 0) then
       Do ich = 1, nch_c; ni = ipsol(ich)
       Do ip = 1, np
        o = woff(ich + nch_c * (ip - 1)); if (o < 0) Cycle
        write(nui) nch_c + ip, ich
        write(nui) w(o + 1:o + ni)
       End do; End do
       Do ip = 1, np; Do jp = 1, ip; k = ibb(ip, jp)
        if (k == 0) Cycle
        if (hpv(k) == 0.d0) Cycle
        write(nui) ip + nch_c, jp + nch_c
        write(nui) hpv(k)
       End do; End do
      end if
      write(nui) 0, 0
      End Subroutine Record_matrix
