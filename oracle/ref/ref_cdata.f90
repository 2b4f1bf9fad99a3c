!======================================================================
! ref_cdata.f90 -- C-callable wrappers around the reference's c_data module
! (DBSR_MAT3/MOD_c_data.f90) with a recording Gen_matrix, so that the list
! order, equal-key merges, block overflow and eps drop of the reference are
! observable from tests.  TEST INFRASTRUCTURE ONLY.
!======================================================================
      Module ref_cdata_log
      Implicit none
      Integer :: nlog = 0, mlog = 0
      Integer, allocatable :: lt(:), lk(:), l1(:), l2(:), l3(:), l4(:), lcall(:)
      Real(8), allocatable :: lc(:)
      Integer :: ncalls = 0
      End Module ref_cdata_log

      Subroutine Gen_matrix(itype, kpol)
      Use c_data
      Use ref_cdata_log
      Implicit none
      Integer, intent(in) :: itype, kpol
      Integer :: j, i
      ncalls = ncalls + 1
      Do j = 1, ncdata
       i = IPT(j)
       if (nlog >= mlog) Stop 'ref_cdata: log full'
       nlog = nlog + 1
       lt(nlog) = itype; lk(nlog) = kpol; lcall(nlog) = ncalls
       l1(nlog) = K1(i); l2(nlog) = K2(i); l3(nlog) = K3(i); l4(nlog) = K4(i)
       lc(nlog) = CDATA(i)
      End do
      End Subroutine Gen_matrix

      Subroutine ref_cdata_alloc(nt, kp1, kp2, mblock, nblock, kblock, eps_c, maxlog) &
        bind(C, name='ref_cdata_alloc')
      Use iso_c_binding
      Use ref_cdata_log
      Implicit none
      Integer(c_int), value :: nt, kp1, kp2, mblock, nblock, kblock, maxlog
      Real(c_double), value :: eps_c
      if (allocated(lt)) Deallocate(lt, lk, l1, l2, l3, l4, lc, lcall)
      mlog = maxlog; nlog = 0; ncalls = 0
      Allocate(lt(mlog), lk(mlog), l1(mlog), l2(mlog), l3(mlog), l4(mlog), lc(mlog), lcall(mlog))
      Call alloc_c_data(nt, kp1, kp2, mblock, nblock, kblock, eps_c)
      End Subroutine ref_cdata_alloc

      Subroutine ref_cdata_add(c, kpol, j1, j2, j3, j4, itype) bind(C, name='ref_cdata_add')
      Use iso_c_binding
      Implicit none
      Real(c_double), value :: c
      Integer(c_int), value :: kpol, j1, j2, j3, j4, itype
      Call Add_coef(c, kpol, j1, j2, j3, j4, itype)
      End Subroutine ref_cdata_add

      Subroutine ref_cdata_flush(itype, kpol) bind(C, name='ref_cdata_flush')
      Use iso_c_binding
      Implicit none
      Integer(c_int), value :: itype, kpol
      Call Add_matrix(itype, kpol)
      End Subroutine ref_cdata_flush

      Function ref_cdata_nlog() result(n) bind(C, name='ref_cdata_nlog')
      Use iso_c_binding
      Use ref_cdata_log
      Implicit none
      Integer(c_int) :: n
      n = nlog
      End Function ref_cdata_nlog

      Subroutine ref_cdata_get(n, ot, ok, o1, o2, o3, o4, oc, ocall) bind(C, name='ref_cdata_get')
      Use iso_c_binding
      Use ref_cdata_log
      Implicit none
      Integer(c_int), value :: n
      Integer(c_int) :: ot(n), ok(n), o1(n), o2(n), o3(n), o4(n), ocall(n)
      Real(c_double) :: oc(n)
      ot = lt(1:n); ok = lk(1:n); o1 = l1(1:n); o2 = l2(1:n); o3 = l3(1:n); o4 = l4(1:n)
      oc = lc(1:n); ocall = lcall(1:n)
      End Subroutine ref_cdata_get
