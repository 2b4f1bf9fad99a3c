!======================================================================
! drv_ci3.f90 -- a host program written the way DBSR_CI3's Setup_matrix
! is (DBSR_CI3/setup_matrix.f90:100-156): for each angular term it calls
! Add_integral(kpol,i,j,j1,j2,j3,j4,CC,mbreit), adds one-electron terms to
! HM/DM itself between them (the met = 1 branch), and ends with
! Add_matrix(jtype,jpol) over all types and multipoles.  Add_integral,
! Add_matrix, Gen_matrix, Rk_evaluate and Sk_evaluate come from
! libdbsr3_amd_ci3.a (dbsr3_amd/fortran/dbs_gpu_ci3.f90).  Reads
! drv_ci3_in.bin, writes drv_ci3_out.bin (tests/test_gpu_fortran.py).
!======================================================================
      Program drv_ci3
      Implicit none
      Integer :: ns, ks, ksp, ksq, norb, ncj, nzero, mbreit, mtype, mpol, mblock, nblock, kblock
      Integer :: nterm, nl, n, l, jtype, jpol
      Real(8), allocatable :: t(:), pq(:,:,:), C(:), lval(:)
      Real(8), allocatable, target :: HM(:,:), DM(:)
      Integer, allocatable :: kef(:), term(:,:), lterm(:,:)

      Open(11, file='drv_ci3_in.bin', access='stream', form='unformatted', status='old')
      Read(11) ns, ks, ksp, ksq
      Allocate(t(ns + ks)); Read(11) t
      Read(11) norb, ncj, nzero, mbreit, mtype, mpol, mblock, nblock, kblock
      Allocate(pq(ns, 2, norb), kef(norb)); Read(11) pq; Read(11) kef
      Read(11) nterm
      Allocate(term(7, nterm), C(nterm)); Read(11) term; Read(11) C
      Read(11) nl
      Allocate(lterm(3, nl), lval(nl)); Read(11) lterm; Read(11) lval
      Close(11)

      Allocate(HM(ncj, nzero), DM(ncj)); HM = 0.d0; DM = 0.d0
      Call dbs_gpu_init(ns, ks, ksp, ksq, t)
      Call dbs_gpu_ci_begin(norb, pq, kef, 2**15, HM, ncj, nzero, DM, 1.d-10, mtype, mpol, mblock, nblock, &
                            kblock, 1.d-12)
      l = 1
      Do n = 1, nterm
! ... one-electron terms due before this angular term (met = 1)
       Do while (l <= nl)
        if (lterm(1, l) /= n) Exit
        if (lterm(3, l) <= nzero) then
         HM(lterm(2, l), lterm(3, l)) = HM(lterm(2, l), lterm(3, l)) + lval(l)
        elseif (lterm(2, l) == lterm(3, l)) then
         DM(lterm(3, l)) = DM(lterm(3, l)) + lval(l)
        end if
        l = l + 1
       End do
       Call Add_integral(term(1, n), term(2, n), term(3, n), term(4, n), term(5, n), term(6, n), term(7, n), &
                         C(n), mbreit)
      End do
! ... final generation of the interaction matrix
      Do jtype = 1, mtype
       Do jpol = 0, mpol
        Call Add_matrix(jtype, jpol)
       End do
      End do
      Call dbs_gpu_ci_end
      Call dbs_gpu_close

      Open(12, file='drv_ci3_out.bin', access='stream', form='unformatted', status='replace')
      Write(12) HM, DM
      Close(12)
      End Program drv_ci3
