!======================================================================
! drv_mat3.f90 -- a host program written the way DBSR_MAT3's State_res is
! (DBSR_MAT3/state_res.f90:160-191): per kind of data (icase 0 overlaps,
! 1 L-integrals, 2 R-integrals, 3 S-integrals) it calls Add_integral(kpol,
! j1,j2,j3,j4,C,ic,jc,io,jo) for every term, then Add_matrix(itype,kpol)
! over all types and multipoles.  Add_integral, Add_matrix, Gen_matrix,
! O_data, L_data, R_data and S_data come from libdbsr3_amd_mat3.a
! (dbsr3_amd/fortran/dbs_gpu_mat3.f90); KBORT and QUADR_PQ are this
! program's own, and it hands over Lval and the one-electron operators, as
! the reference's host programs hold them.  Reads drv_mat3_in.bin, writes
! drv_mat3_out.bin (tests/test_gpu_fortran.py).
!======================================================================
      Module drv_mat3_data
      Implicit none
      Integer :: norb
      Integer, allocatable :: kb(:,:)
      Real(8), allocatable :: pq(:,:,:)
      End Module drv_mat3_data

      Integer Function KBORT(i1, i2)
      Use drv_mat3_data
      Implicit none
      Integer, intent(in) :: i1, i2
      KBORT = kb(i1, i2)
      End Function KBORT

      Real(8) Function QUADR_PQ(i1, i2, m)
      Use drv_mat3_data
      Implicit none
      Integer, intent(in) :: i1, i2, m
      Real(8), external :: quadr
      QUADR_PQ = quadr(pq(1, 1, i1), pq(1, 1, i2), m)
      End Function QUADR_PQ

      Program drv_mat3
      Use drv_mat3_data
      Implicit none
      Integer :: ns, ks, ksp, ksq, nbound, ibo, mk, nch, npert, nvch, iitar, nterm, ntarg, n, icase, it, kp
      Integer :: nblk, ncb, nbb, nkap
      Real(8), allocatable :: lval(:,:), hd(:,:,:)
      Integer, allocatable :: kap_list(:)
      Real(8), allocatable :: t(:), vch(:,:), fppqq(:,:), C(:), htarg(:), otarg(:), hch(:), acf(:), hcp(:), hp(:)
      Integer, allocatable :: lbs(:), kbs(:), ipbs(:), icc(:,:), icb(:,:), ibb(:,:), iptar(:), term(:,:)

      Open(11, file='drv_mat3_in.bin', access='stream', form='unformatted', status='old')
      Read(11) ns, ks, ksp, ksq
      Allocate(t(ns + ks)); Read(11) t
      Read(11) norb, nbound, ibo, mk, nch, npert, nvch, iitar, ntarg
      Allocate(pq(ns, 2, norb), lbs(norb), kbs(norb), ipbs(norb), kb(norb, norb))
      Read(11) pq; Read(11) lbs; Read(11) kbs; Read(11) ipbs; Read(11) kb
      Allocate(icc(nch, nch), icb(nch, max(npert, 1)), ibb(max(npert, 1), max(npert, 1)), iptar(nch))
      Read(11) icc; Read(11) icb; Read(11) ibb; Read(11) iptar
      Allocate(vch(2 * ns, nvch), fppqq(2 * ns, 2 * ns)); Read(11) vch; Read(11) fppqq
      Read(11) nterm
      Allocate(term(10, nterm), C(nterm)); Read(11) term; Read(11) C
      Read(11) nkap
      Allocate(lval(norb, norb), kap_list(nkap), hd(2 * ns, 2 * ns, nkap))
      Read(11) lval; Read(11) kap_list; Read(11) hd
      Close(11)

      Allocate(htarg(ntarg), otarg(ntarg)); htarg = 0.d0; otarg = 0.d0
      Call dbs_gpu_init(ns, ks, ksp, ksq, t)
      Call dbs_gpu_mat_begin(norb, nbound, pq, lbs, kbs, ipbs, ibo, mk, 1.d-10, 1, nch, npert, icc, icb, ibb, &
                             nvch, vch, fppqq, iitar, iptar, htarg, otarg, 19, 64, 400, 40, 1.d-14)
      Call dbs_gpu_mat_one_electron(lval, nkap, kap_list, hd)
      Do icase = 0, 3
       Call dbs_gpu_mat_case(icase)
       Do n = 1, nterm
        if (term(1, n) /= icase) Cycle
        Call Add_integral(term(2, n), term(3, n), term(4, n), term(5, n), term(6, n), C(n), term(7, n), &
                          term(8, n), term(9, n), term(10, n))
       End do
       Do kp = 0, mk + 1
        Do it = 1, 19
         Call Add_matrix(it, kp)
        End do
       End do
      End do
      nblk = maxval(icc); ncb = maxval(icb); nbb = maxval(ibb)
      Allocate(hch(4 * ns * ns * nblk), acf(nblk * (mk + 1)), hcp(2 * ns * max(ncb, 1)), hp(max(nbb, 1)))
      Call dbs_gpu_mat_fetch(hch, acf, hcp, hp)
      Call dbs_gpu_mat_end
      Call dbs_gpu_close

      Open(12, file='drv_mat3_out.bin', access='stream', form='unformatted', status='replace')
      Write(12) hch, acf, hcp(1:2 * ns * ncb), hp(1:nbb), htarg, otarg
      Close(12)
      End Program drv_mat3
