!======================================================================
! dbs_gpu_mat3.f90 -- DBSR_MAT3's accumulation path on the GPU with the
! reference's names and argument lists:
!     Call Add_integral(kpol,i1,i2,i3,i4,C,ic,jc,io,jo)
!                                     (DBSR_MAT3/add_integral.f90:2-57)
!     Call Add_matrix(jtype,jpol)     (DBSR_MAT3/MOD_c_data.f90:318-368)
!     Call Gen_matrix(itype,kpol)     (DBSR_MAT3/gen_matrix.f90:2-17)
!     Call O_data(itype)              (DBSR_MAT3/O_data.f90:2-93)
!     Call L_data(itype)              (DBSR_MAT3/L_data.f90:2-145)
!     Call R_data(itype,kpol,btype)   (DBSR_MAT3/R_data.f90:2-290)
!     Call S_data(ktype,kpol)         (DBSR_MAT3/S_data.f90:2-287)
! and (dbs_gpu_mat3_post.f90) Zero_cond, DBS_ORTH, Diag_channels and
! Record_matrix.
! Module dbs_gpu_mat plays the part of module dbsr_mat (icase, ibo, mk,
! eps_c, check_target, lbs/kbs/ipbs, V_ch, fppqq, iitar, iptar and the
! matrices hch, acf, hcp, hp with their tables icc, icb, ibb): the host
! hands it over with dbs_gpu_mat_begin, sets icase with dbs_gpu_mat_case
! as State_res does, and takes the matrices back with dbs_gpu_mat_fetch.
! The host program's own KBORT(i1,i2) and QUADR_PQ(i1,i2,m) (Libraries_03)
! are called back through bind(C) trampolines.  For L-integrals (icase 1)
! the host hands over, with dbs_gpu_mat_one_electron, what L_data reads
! from Libraries_03: Lval(i,j) of the orbitals as a table and the
! one-electron operators hd(ms,ms,nkap) of module dhl_core.  For the
! channel post-processing the host hands over, with dbs_gpu_mat_channels,
! the grid sizes of DBS_grid, the channels' l, kappa and orbital
! (channel_jj), the zero conditions, the energy window, idiag and the unit
! Record_matrix writes; the diagonal blocks' overlaps are kept on the
! device by dbs_gpu_mat_save_diag (sub1's diag = hch copy), and the host's
! own IBORT(i,j), Get_pv(j,v,ns) and Get_qv(j,v,ns) are called for the
! orthogonality conditions.  Built into libdbsr3_amd_mat3.a, linked into
! the host program.
!======================================================================
      Module dbs_gpu_mat
      Use iso_c_binding
      Implicit none
      Type, bind(C) :: mat_blocks_t        ! struct dbs_mat_blocks
       Integer(c_int) :: nch, npert, mk
       Type(c_ptr) :: icc
       Integer(c_int) :: nblocks
       Type(c_ptr) :: hch_dev, acf_dev, icb
       Integer(c_int) :: ncb
       Type(c_ptr) :: hcp_dev, ibb
       Integer(c_int) :: nbb
       Type(c_ptr) :: hp_dev
      End Type
      Type, bind(C) :: mat3_aux_t          ! struct dbs_mat3_aux
       Integer(c_int) :: ibo
       Type(c_ptr) :: ipbs
       Type(c_funptr) :: kbort
       Type(c_ptr) :: kbort_user, vch_dev
       Integer(c_int) :: nvch
       Type(c_ptr) :: fppqq_dev
       Integer(c_int) :: iitar
       Type(c_ptr) :: iptar, htarg
       Type(c_funptr) :: quadr_pq
       Type(c_ptr) :: quadr_pq_user
      End Type
      Type, bind(C) :: mat3_orbs_t         ! struct dbs_mat3_orbs
       Integer(c_int) :: norb
       Type(c_ptr) :: lbs, kbs, ipbs
       Integer(c_int) :: ibo, mk
       Real(c_double) :: eps_c
       Integer(c_int) :: check_target
      End Type
      Type, bind(C) :: ldata_aux_t         ! struct dbs_ldata_aux
       Type(c_ptr) :: mat3, otarg
       Integer(c_int) :: check_target, norb
       Type(c_ptr) :: lval, kbs
       Integer(c_int) :: nkap
       Type(c_ptr) :: kap_list, hd_dev
      End Type
      Type(mat_blocks_t), target :: mb
      Type(mat3_aux_t), target :: aux
      Type(mat3_orbs_t), target :: orbs
      Type(c_ptr) :: cdata = c_null_ptr, pq_dev = c_null_ptr
      Integer :: icase = 2, norb = 0, nbound = 0, ms = 0
      Real(8) :: eps_c = 1.d-10
      Integer(c_int), allocatable, target :: lbs(:), kbs(:), ipbs(:), ipbs0(:), icc(:,:), icb(:,:), &
                                             ibb(:,:), iptar(:)
      Type(c_ptr) :: otarg = c_null_ptr
      Integer :: check_t = 1
      ! L_data's one-electron data (dbs_gpu_mat_one_electron)
      Type(ldata_aux_t), target :: laux
      Real(8), allocatable, target :: lval(:,:)
      Integer(c_int), allocatable, target :: kap_list(:)
      Type(c_ptr) :: hd_dev = c_null_ptr
      Integer :: nc_cur = 0
      Type(c_ptr) :: c_cur, k1_cur, k2_cur, k3_cur, k4_cur
      ! channel post-processing (dbs_gpu_mat_channels, dbs_gpu_mat_save_diag)
      Integer :: idiag = 0, nui = 0, ilzero = -1, jlzero = -1, ibzero = 1, jbzero = 1
      Integer :: nsp_c = 0, nsq_c = 0, ksp_c = 0, ksq_c = 0, nch_c = 0, nhm = 0
      Real(8) :: Edmin = -2.d0 * 137.03599976d0**2, Edmax = 1.d10, Egap = 1.d-3
      Integer(c_int), allocatable, target :: lch(:), kch(:), ipch(:), iprm(:), ipsol(:)
      Real(8), allocatable, target :: overlaps(:,:)
      Type(c_ptr) :: diag_dev = c_null_ptr
      End Module dbs_gpu_mat

!----------------------------------------------------------------------
!     trampolines to the host program's KBORT and QUADR_PQ
!----------------------------------------------------------------------
      Function dbs_gpu_kbort(user, i1, i2) bind(C) result(k)
      Use iso_c_binding
      Implicit none
      Type(c_ptr), value :: user
      Integer(c_int), value :: i1, i2
      Integer(c_int) :: k
      Integer, external :: KBORT
      k = KBORT(i1, i2)
      End Function dbs_gpu_kbort

      Function dbs_gpu_quadr_pq(user, i1, i2, m) bind(C) result(s)
      Use iso_c_binding
      Implicit none
      Type(c_ptr), value :: user
      Integer(c_int), value :: i1, i2, m
      Real(c_double) :: s
      Real(8), external :: QUADR_PQ
      s = QUADR_PQ(i1, i2, m)
      End Function dbs_gpu_quadr_pq

      Subroutine dbs_gpu_mat_gen(user, itype, kpol, nc, c, K1, K2, K3, K4) bind(C)
      Use dbs_gpu_mat
      Implicit none
      Type(c_ptr), value :: user, c, K1, K2, K3, K4
      Integer(c_int), value :: itype, kpol, nc
      nc_cur = nc; c_cur = c; k1_cur = K1; k2_cur = K2; k3_cur = K3; k4_cur = K4
      Call Gen_matrix(itype, kpol)
      nc_cur = 0
      End Subroutine dbs_gpu_mat_gen

!----------------------------------------------------------------------
      Subroutine dbs_gpu_mat_begin(norb_, nbound_, pq, lbs_, kbs_, ipbs_, ibo_, mk_, eps_c_, check_target_, &
                                   nch, npert, icc_, icb_, ibb_, nvch, vch, fppqq, iitar_, iptar_, htarg, &
                                   otarg_, ntype, mblock, nblock, kblock, eps_cdata)
!----------------------------------------------------------------------
!     module dbsr_mat handed over: orbital tables lbs/kbs/ipbs(1:norb),
!     bound-orbital functions pq(ns,2,nbound) (the orbitals the lists
!     name, 1..nbound), ibo, mk, eps_c, check_target; the storage tables
!     icc(nch,nch), icb(nch,npert), ibb(npert,npert) of alloc_dbsr_matrix
!     (the matrices themselves are allocated on the device, zero);
!     V_ch(ms,nvch), fppqq(ms,ms); iitar, iptar(nch); the packed target
!     interactions htarg and overlaps otarg (host arrays updated in place);
!     then the coefficient lists of types 1..ntype, multipoles 0..mk
      Use dbs_gpu_api
      Use dbs_gpu_mem
      Use dbs_gpu_state, only: ctx, ns_g
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(in) :: norb_, nbound_, ibo_, mk_, check_target_, nch, npert, nvch, iitar_, ntype, &
                             mblock, nblock, kblock
      Real(8), intent(in), target :: pq(ns_g, 2, nbound_), vch(2 * ns_g, *), fppqq(2 * ns_g, 2 * ns_g)
      Integer, intent(in) :: lbs_(norb_), kbs_(norb_), ipbs_(norb_), icc_(nch, nch), icb_(nch, *), &
                             ibb_(npert, *), iptar_(nch)
      Real(8), intent(inout), target :: htarg(*), otarg_(*)
      Real(8), intent(in) :: eps_c_, eps_cdata
      Integer :: nblk, ncb, nbb
      Interface
       Function dbs_gpu_kbort(user, i1, i2) bind(C) result(k)
        Import :: c_ptr, c_int
        Type(c_ptr), value :: user
        Integer(c_int), value :: i1, i2
        Integer(c_int) :: k
       End Function
       Function dbs_gpu_quadr_pq(user, i1, i2, m) bind(C) result(s)
        Import :: c_ptr, c_int, c_double
        Type(c_ptr), value :: user
        Integer(c_int), value :: i1, i2, m
        Real(c_double) :: s
       End Function
      Subroutine dbs_gpu_mat_gen(user, itype, kpol, nc, c, K1, K2, K3, K4) bind(C)
        Import :: c_ptr, c_int
        Type(c_ptr), value :: user, c, K1, K2, K3, K4
        Integer(c_int), value :: itype, kpol, nc
       End Subroutine
      End Interface
      if (.not. c_associated(ctx)) Stop 'dbs_gpu: call dbs_gpu_init first'
      Call dbs_gpu_mat_end
      norb = norb_; nbound = nbound_; eps_c = eps_c_; ms = 2 * ns_g
      Allocate(lbs(norb), kbs(norb), ipbs(norb), ipbs0(0:norb), icc(nch, nch), icb(nch, max(npert, 1)), &
               ibb(max(npert, 1), max(npert, 1)), iptar(0:nch))
      lbs = lbs_; kbs = kbs_; ipbs = ipbs_; ipbs0(0) = 0; ipbs0(1:norb) = ipbs_
      icc = icc_; icb = 0; ibb = 0; iptar(0) = 0; iptar(1:nch) = iptar_
      if (npert > 0) then
       icb(:, 1:npert) = icb_(:, 1:npert); ibb(1:npert, 1:npert) = ibb_(1:npert, 1:npert)
      end if
      nblk = maxval(icc); ncb = maxval(icb); nbb = maxval(ibb)
      mb%nch = nch; mb%npert = npert; mb%mk = mk_
      mb%icc = c_loc(icc); mb%nblocks = nblk; mb%icb = c_loc(icb); mb%ncb = ncb
      mb%ibb = c_loc(ibb); mb%nbb = nbb
      Call dbs_gpu_zeros(mb%hch_dev, ms * ms * max(nblk, 1))
      Call dbs_gpu_zeros(mb%acf_dev, max(nblk, 1) * (mk_ + 1))
      Call dbs_gpu_zeros(mb%hcp_dev, ms * max(ncb, 1))
      Call dbs_gpu_zeros(mb%hp_dev, max(nbb, 1))
      pq_dev = dbs_gpu_put(pq, ns_g * 2 * nbound)
      aux%ibo = ibo_; aux%ipbs = c_loc(ipbs0(1)); aux%kbort = c_funloc(dbs_gpu_kbort)
      aux%kbort_user = c_null_ptr; aux%vch_dev = dbs_gpu_put(vch, ms * max(nvch, 1)); aux%nvch = nvch
      aux%fppqq_dev = dbs_gpu_put(fppqq, ms * ms); aux%iitar = iitar_; aux%iptar = c_loc(iptar(1))
      aux%htarg = c_loc(htarg); aux%quadr_pq = c_funloc(dbs_gpu_quadr_pq); aux%quadr_pq_user = c_null_ptr
      otarg = c_loc(otarg_)
      orbs%norb = norb; orbs%lbs = c_loc(lbs); orbs%kbs = c_loc(kbs); orbs%ipbs = c_loc(ipbs)
      orbs%ibo = ibo_; orbs%mk = mk_; orbs%eps_c = eps_c_; orbs%check_target = check_target_
      check_t = check_target_
      Call dbs_gpu_check(dbs_cdata_create(cdata, ntype, 0, mk_ + 1, mblock, nblock, kblock, eps_cdata, &
                         c_funloc(dbs_gpu_mat_gen), c_null_ptr))
      End Subroutine dbs_gpu_mat_begin

      Subroutine dbs_gpu_zeros(p, n)
      Use dbs_gpu_api
      Use dbs_gpu_mem
      Implicit none
      Type(c_ptr), intent(out) :: p
      Integer, intent(in) :: n
      Real(8), allocatable, target :: z(:)
      Allocate(z(n)); z = 0.d0
      p = dbs_gpu_put(z, n)
      Deallocate(z)
      End Subroutine dbs_gpu_zeros

      Subroutine dbs_gpu_mat_one_electron(lval_, nkap, kap_list_, hd)
!     L_data's inputs from Libraries_03: Lval(i,j) of the orbitals 1..norb
!     and the one-electron operators hd(ms,ms,nkap) of kappas kap_list
      Use dbs_gpu_api
      Use dbs_gpu_mem
      Use dbs_gpu_state, only: ctx, ns_g
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(in) :: nkap, kap_list_(nkap)
      Real(8), intent(in) :: lval_(norb, norb)
      Real(8), intent(in), target :: hd(2 * ns_g, 2 * ns_g, nkap)
      if (.not. c_associated(cdata)) Stop 'dbs_gpu_mat_one_electron: call dbs_gpu_mat_begin first'
      if (allocated(lval)) Deallocate(lval, kap_list)
      if (c_associated(hd_dev)) Call dbs_gpu_check(dbs_free(ctx, hd_dev))
      Allocate(lval(norb, norb), kap_list(nkap))
      lval = lval_; kap_list = kap_list_
      hd_dev = dbs_gpu_put(hd, ms * ms * nkap)
      laux%mat3 = c_loc(aux); laux%otarg = otarg; laux%check_target = check_t; laux%norb = norb
      laux%lval = c_loc(lval); laux%kbs = c_loc(kbs); laux%nkap = nkap; laux%kap_list = c_loc(kap_list)
      laux%hd_dev = hd_dev
      End Subroutine dbs_gpu_mat_one_electron

      Subroutine dbs_gpu_mat_case(icase_)
!     the kind of data State_res feeds next: 0 overlaps, 1 L, 2 R, 3 S
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(in) :: icase_
      icase = icase_
      End Subroutine dbs_gpu_mat_case

      Subroutine dbs_gpu_mat_fetch(hch, acf, hcp, hp)
!     the matrices back to the host: hch(ms,ms,iicc), acf(iicc,0:mk),
!     hcp(ms,iicb), hp(iibb)
      Use dbs_gpu_api
      Use dbs_gpu_mem
      Use dbs_gpu_mat
      Implicit none
      Real(8), intent(out) :: hch(*), acf(*), hcp(*), hp(*)
      Call dbs_gpu_get(mb%hch_dev, hch, ms * ms * mb%nblocks)
      Call dbs_gpu_get(mb%acf_dev, acf, mb%nblocks * (mb%mk + 1))
      Call dbs_gpu_get(mb%hcp_dev, hcp, ms * mb%ncb)
      Call dbs_gpu_get(mb%hp_dev, hp, mb%nbb)
      End Subroutine dbs_gpu_mat_fetch

      Subroutine dbs_gpu_mat_put(hch, acf, hcp, hp)
!     matrices from the host into the current storage (the inverse of
!     dbs_gpu_mat_fetch)
      Use dbs_gpu_api
      Use dbs_gpu_mem
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_mat
      Implicit none
      Real(8), intent(in), target :: hch(*), acf(*), hcp(*), hp(*)
      if (mb%nblocks > 0) Call dbs_gpu_check(dbs_memcpy_h2d(ctx, mb%hch_dev, c_loc(hch), &
                                             int(8, c_size_t) * ms * ms * mb%nblocks))
      if (mb%nblocks > 0) Call dbs_gpu_check(dbs_memcpy_h2d(ctx, mb%acf_dev, c_loc(acf), &
                                             int(8, c_size_t) * mb%nblocks * (mb%mk + 1)))
      if (mb%ncb > 0) Call dbs_gpu_check(dbs_memcpy_h2d(ctx, mb%hcp_dev, c_loc(hcp), int(8, c_size_t) * ms * mb%ncb))
      if (mb%nbb > 0) Call dbs_gpu_check(dbs_memcpy_h2d(ctx, mb%hp_dev, c_loc(hp), int(8, c_size_t) * mb%nbb))
      End Subroutine dbs_gpu_mat_put

      Subroutine dbs_gpu_mat_channels(nsp, nsq, ksp, ksq, lch_, kch_, ipch_, ilzero_, jlzero_, ibzero_, jbzero_, &
                                      edmin_, edmax_, egap_, idiag_, nui_)
!     the post-processing's data of modules DBS_grid (nsp, nsq, ksp, ksq),
!     channel_jj (l, kappa and orbital of each channel) and dbsr_mat (zero
!     conditions, energy window, idiag of the current storage, unit nui)
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(in) :: nsp, nsq, ksp, ksq, lch_(*), kch_(*), ipch_(*), ilzero_, jlzero_, ibzero_, jbzero_, &
                             idiag_, nui_
      Real(8), intent(in) :: edmin_, edmax_, egap_
      if (.not. c_associated(cdata)) Stop 'dbs_gpu_mat_channels: call dbs_gpu_mat_begin first'
      nch_c = mb%nch
      if (allocated(lch)) Deallocate(lch, kch, ipch)
      Allocate(lch(max(nch_c, 1)), kch(max(nch_c, 1)), ipch(max(nch_c, 1)))
      lch(1:nch_c) = lch_(1:nch_c); kch(1:nch_c) = kch_(1:nch_c); ipch(1:nch_c) = ipch_(1:nch_c)
      nsp_c = nsp; nsq_c = nsq; ksp_c = ksp; ksq_c = ksq
      ilzero = ilzero_; jlzero = jlzero_; ibzero = ibzero_; jbzero = jbzero_
      Edmin = edmin_; Edmax = edmax_; Egap = egap_; idiag = idiag_; nui = nui_
      End Subroutine dbs_gpu_mat_channels

      Subroutine dbs_gpu_mat_save_diag
!     diag(:,:,ich) = hch(:,:,icc(ich,ich)), kept on the device for
!     Diag_channels and Record_matrix (DBSR_MAT3/sub1.f90:98-99)
      Use dbs_gpu_api
      Use dbs_gpu_mem
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_mat
      Implicit none
      Real(8), allocatable, target :: blk(:)
      Integer :: ich, k
      Integer(c_size_t) :: nb
      if (.not. c_associated(cdata)) Stop 'dbs_gpu_mat_save_diag: call dbs_gpu_mat_begin first'
      nb = int(8, c_size_t) * ms * ms
      if (c_associated(diag_dev)) Call dbs_gpu_check(dbs_free(ctx, diag_dev))
      Call dbs_gpu_check(dbs_malloc(ctx, diag_dev, nb * max(mb%nch, 1)))
      Allocate(blk(ms * ms))
      Do ich = 1, mb%nch
       k = icc(ich, ich); blk = 0.d0
       if (k > 0) Call dbs_gpu_check(dbs_memcpy_d2h(ctx, c_loc(blk), &
                                     transfer(transfer(mb%hch_dev, 0_c_intptr_t) + nb * (k - 1), c_null_ptr), nb))
       Call dbs_gpu_check(dbs_memcpy_h2d(ctx, transfer(transfer(diag_dev, 0_c_intptr_t) + nb * (ich - 1), &
                          c_null_ptr), c_loc(blk), nb))
      End do
      End Subroutine dbs_gpu_mat_save_diag

      Subroutine dbs_gpu_mat_fetch_channels(iprm_, nhm_, ipsol_, diag)
!     Zero_cond's iprm(nch*ms+npert) and nhm, Diag_channels' ipsol(0:nch)
!     and diag(ms,ms,nch) back to the host
      Use dbs_gpu_mem
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(out) :: iprm_(*), nhm_, ipsol_(0:*)
      Real(8), intent(out) :: diag(*)
      if (.not. allocated(ipsol)) Stop 'dbs_gpu_mat_fetch_channels: call Diag_channels first'
      iprm_(1:size(iprm)) = iprm; nhm_ = nhm; ipsol_(0:nch_c) = ipsol(0:nch_c)
      Call dbs_gpu_get(diag_dev, diag, ms * ms * nch_c)
      End Subroutine dbs_gpu_mat_fetch_channels

      Subroutine dbs_gpu_mat_overlaps(ov)
!     the overlaps(nch+npert,nch+npert) of the last Record_matrix with icase 0
      Use dbs_gpu_mat
      Implicit none
      Real(8), intent(out) :: ov(*)
      Integer :: n
      if (.not. allocated(overlaps)) Stop 'dbs_gpu_mat_overlaps: no Record_matrix with icase 0'
      n = size(overlaps, 1)
      ov(1:n * n) = reshape(overlaps, [n * n])
      End Subroutine dbs_gpu_mat_overlaps

      Subroutine dbs_gpu_mat_channels_end
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_mat
      Implicit none
      if (c_associated(diag_dev)) Call dbs_gpu_check(dbs_free(ctx, diag_dev))
      diag_dev = c_null_ptr
      if (allocated(lch)) Deallocate(lch, kch, ipch)
      if (allocated(iprm)) Deallocate(iprm)
      if (allocated(ipsol)) Deallocate(ipsol)
      if (allocated(overlaps)) Deallocate(overlaps)
      End Subroutine dbs_gpu_mat_channels_end

      Subroutine dbs_gpu_mat_end
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_mat
      Implicit none
      if (c_associated(cdata)) then
       Call dbs_gpu_check(dbs_cdata_destroy(cdata))
       Call dbs_gpu_check(dbs_free(ctx, mb%hch_dev)); Call dbs_gpu_check(dbs_free(ctx, mb%acf_dev))
       Call dbs_gpu_check(dbs_free(ctx, mb%hcp_dev)); Call dbs_gpu_check(dbs_free(ctx, mb%hp_dev))
       Call dbs_gpu_check(dbs_free(ctx, pq_dev)); Call dbs_gpu_check(dbs_free(ctx, aux%vch_dev))
       Call dbs_gpu_check(dbs_free(ctx, aux%fppqq_dev))
      end if
      cdata = c_null_ptr
      if (allocated(lbs)) Deallocate(lbs, kbs, ipbs, ipbs0, icc, icb, ibb, iptar)
      if (allocated(lval)) Deallocate(lval, kap_list)
      if (c_associated(hd_dev)) Call dbs_gpu_check(dbs_free(ctx, hd_dev))
      hd_dev = c_null_ptr
      End Subroutine dbs_gpu_mat_end

!======================================================================
      Subroutine Add_integral(kpol, i1, i2, i3, i4, C, ic, jc, io, jo)
!======================================================================
!     DBSR_MAT3/add_integral.f90:2-57: icase 0 -> Idef_Otype, 1 ->
!     Idef_Ltype, 2 -> parity rule and Idef_Rtype, 3 -> SMU and Idef_itype,
!     into the lists
      Use dbs_gpu_api
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(in) :: kpol, i1, i2, i3, i4, ic, jc, io, jo
      Real(8), intent(in) :: C
      Call dbs_gpu_check(dbs_add_integral_mat3(cdata, c_loc(orbs), icase, kpol, i1, i2, i3, i4, C, ic, jc, &
                         io, jo))
      End Subroutine Add_integral

      Subroutine Add_matrix(jtype, jpol)
      Use dbs_gpu_api
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(in) :: jtype, jpol
      Call dbs_gpu_check(dbs_cdata_add_matrix(cdata, jtype, jpol))
      End Subroutine Add_matrix

!======================================================================
      Subroutine Gen_matrix(itype, kpol)
!======================================================================
!     DBSR_MAT3/gen_matrix.f90:2-17
      Use dbs_gpu_mat, only: icase
      Implicit none
      Integer, intent(in) :: itype, kpol
      Select case(icase)
       Case(0); Call O_data(itype)
       Case(1); Call L_data(itype)
       Case(2)
        Call R_data(itype, kpol, 0)
        Call R_data(itype, kpol, 1)
        Call R_data(itype, kpol, 2)
        Call R_data(itype, kpol, 3)
       Case(3); Call S_data(itype, kpol)
      End Select
      End Subroutine Gen_matrix

      Subroutine O_data(itype)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(in) :: itype
      Interface
       Function dbs_gpu_kbort(user, i1, i2) bind(C) result(k)
        Import :: c_ptr, c_int
        Type(c_ptr), value :: user
        Integer(c_int), value :: i1, i2
        Integer(c_int) :: k
       End Function
      End Interface
      if (nc_cur <= 0) Return
      Call dbs_gpu_check(dbs_o_data_dev(ctx, c_loc(mb), itype, nc_cur, k1_cur, k2_cur, c_cur, eps_c, aux%ibo, &
                         aux%ipbs, c_funloc(dbs_gpu_kbort), c_null_ptr, aux%vch_dev, aux%nvch, aux%fppqq_dev, &
                         aux%iitar, aux%iptar, otarg))
      End Subroutine O_data

      Subroutine L_data(itype)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(in) :: itype
      if (nc_cur <= 0) Return
      if (.not. allocated(lval)) Stop 'L_data: call dbs_gpu_mat_one_electron first'
      Call dbs_gpu_check(dbs_ldata_dev(ctx, c_loc(mb), c_loc(laux), itype, nc_cur, k1_cur, k2_cur, k3_cur, &
                         k4_cur, c_cur, eps_c, pq_dev))
      End Subroutine L_data

      Subroutine R_data(itype, kpol, btype)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(in) :: itype, kpol, btype
      if (nc_cur <= 0) Return
      Call dbs_gpu_check(dbs_rdata_dev(ctx, c_loc(mb), c_loc(aux), itype, btype, kpol, nc_cur, c_cur, k1_cur, &
                         k2_cur, k3_cur, k4_cur, pq_dev, nbound))
      End Subroutine R_data

      Subroutine S_data(ktype, kpol)
      Use dbs_gpu_api
      Use dbs_gpu_state, only: ctx
      Use dbs_gpu_mat
      Implicit none
      Integer, intent(in) :: ktype, kpol
      if (nc_cur <= 0) Return
      Call dbs_gpu_check(dbs_sdata_dev(ctx, c_loc(mb), c_loc(aux), ktype, kpol, nc_cur, c_cur, k1_cur, k2_cur, &
                         k3_cur, k4_cur, pq_dev, nbound))
      End Subroutine S_data

