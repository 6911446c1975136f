!-----------------------------------------------------------------------
! Fortran caller of the two-density drop-ins with every per-density array
! allocated on its own, as the reference allocates them:
!   UHF CHO_FSCF and CHO_LK_SCF with nDen = 2 (FLT(1:2), Cka(1:2), KLT(1:2),
!     src/scf/choscf_drv.F90:285-291,401-402,420,433),
!   CHO_FMCSCF and CHO_LK_CASSCF with DoActive (FI/FA and DI/DA,
!     src/fock_util/cho_cas_drv.F90:42-43,101-102,277,285).
! Every array is followed by nCan canary words; outputs are checked after
! each call, inputs carry huge values there so that a read past the end
! shows in the results.  Reads from a raw FP64 file
!   nBas, NumCho, nOa, nOb, nI, nA, nChM, nFro, nOrb, ExFac, FactXI,
!   L(nP,NumCho), PorbA(nOa,nBas), PorbB(nOb,nBas), MSQa(nBas,nBas),
!   MSQb(nBas,nBas), DLT, FLTa0, FLTb0, DI, DA, FI0, FA0,
!   POrb1(nI,nBas), POrb2(nChM,nBas), POrb3(nA,nBas), CMO(nBas,nBas),
!   MSQi(nBas,nBas), Ash2(nBas,nBas)
! and writes FLTa, FLTb (FSCF), FLTa, FLTb, KLTa, KLTb (LK_SCF), FI, FA, W
! (FMCSCF), FI, FA, W (LK_CASSCF) and the number of damaged canaries.
! Usage: uhf_cas_driver <input.bin> <output.bin>
!-----------------------------------------------------------------------
program uhf_cas_driver
  use, intrinsic :: iso_c_binding, only: c_int64_t, c_int32_t, c_double
  use molcas_jk_iface
  implicit none
  integer, parameter :: nCan = 64
  real(8), parameter :: canary = 1.0e300_8
  character(len=512) :: fin, fout
  integer :: u, rc, nBas, NumCho, nOa, nOb, nI, nA, nChM, nFro, nOrb, nP, nW, i, nbad
  real(8) :: hdr(11), ExFac, FactXI
  real(8), allocatable :: L(:), PorbA(:), PorbB(:), MSQa(:), MSQb(:), DLT(:), FLTa0(:), FLTb0(:)
  real(8), allocatable :: DI(:), DA(:), FI0(:), FA0(:), POrb1(:), POrb2(:), POrb3(:), CMO(:), MSQi(:), Ash2(:)
  real(8), allocatable :: FLTa(:), FLTb(:), LKa(:), LKb(:), KLTa(:), KLTb(:)
  real(8), allocatable :: FI(:), FA(:), W(:), FI2(:), FA2(:), W2(:)
  integer(c_int64_t), allocatable :: map(:)
  integer :: nForb(8, 2), nIorb(8, 2), nBasSh(1)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  nBas = int(hdr(1)); NumCho = int(hdr(2)); nOa = int(hdr(3)); nOb = int(hdr(4))
  nI = int(hdr(5)); nA = int(hdr(6)); nChM = int(hdr(7)); nFro = int(hdr(8)); nOrb = int(hdr(9))
  ExFac = hdr(10); FactXI = hdr(11)
  nP = nBas*(nBas+1)/2
  nW = max(1, nOrb*nA*(nA*(nA+1)/2))
  ! inputs: the payload, then nCan huge words
  call alloc_in(L, nP*NumCho)
  call alloc_in(PorbA, nOa*nBas)
  call alloc_in(PorbB, nOb*nBas)
  call alloc_in(MSQa, nBas*nBas)
  call alloc_in(MSQb, nBas*nBas)
  call alloc_in(DLT, nP)
  call alloc_in(FLTa0, nP)
  call alloc_in(FLTb0, nP)
  call alloc_in(DI, nP)
  call alloc_in(DA, nP)
  call alloc_in(FI0, nP)
  call alloc_in(FA0, nP)
  call alloc_in(POrb1, nI*nBas)
  call alloc_in(POrb2, nChM*nBas)
  call alloc_in(POrb3, nA*nBas)
  call alloc_in(CMO, nBas*nBas)
  call alloc_in(MSQi, nBas*nBas)
  call alloc_in(Ash2, nBas*nBas)
  read(u) L(1:nP*NumCho)
  read(u) PorbA(1:nOa*nBas)
  read(u) PorbB(1:nOb*nBas)
  read(u) MSQa(1:nBas*nBas)
  read(u) MSQb(1:nBas*nBas)
  read(u) DLT(1:nP)
  read(u) FLTa0(1:nP)
  read(u) FLTb0(1:nP)
  read(u) DI(1:nP)
  read(u) DA(1:nP)
  read(u) FI0(1:nP)
  read(u) FA0(1:nP)
  read(u) POrb1(1:nI*nBas)
  read(u) POrb2(1:nChM*nBas)
  read(u) POrb3(1:nA*nBas)
  read(u) CMO(1:nBas*nBas)
  read(u) MSQi(1:nBas*nBas)
  read(u) Ash2(1:nBas*nBas)
  close(u)
  ! outputs: separately allocated, canaries after the payload
  call alloc_out(FLTa, nP, FLTa0)
  call alloc_out(FLTb, nP, FLTb0)
  call alloc_out(LKa, nP, FLTa0)
  call alloc_out(LKb, nP, FLTb0)
  call alloc_out(KLTa, nP)
  call alloc_out(KLTb, nP)
  call alloc_out(FI, nP, FI0)
  call alloc_out(FA, nP, FA0)
  call alloc_out(W, nW)
  call alloc_out(FI2, nP, FI0)
  call alloc_out(FA2, nP, FA0)
  call alloc_out(W2, nW)
  allocate(map(nP))
  do i = 1, nP
    map(i) = int(i-1, c_int64_t)
  end do

  call jk_gpu_init(rc, nBas, NumCho)
  if (rc /= 0) stop 11
  rc = int(mjk_set_reduced_set(jk_ctx, 0_c_int32_t, int(nP, c_int64_t), map))
  if (rc /= 0) stop 12
  rc = int(mjk_upload_vectors(jk_ctx, 0_c_int32_t, 0_c_int64_t, int(NumCho, c_int64_t), L, 0_c_int32_t))
  if (rc /= 0) stop 13
  nForb = 0
  nIorb = 0
  nIorb(1, 1) = nOa
  nIorb(1, 2) = nOb
  ! UHF: src/scf/choscf_drv.F90:420 (ALGO 3) and :433 (ALGO 4)
  call cho_fscf_gpu(rc, 2, FLTa, FLTb, nForb, nIorb, PorbA, PorbB, DLT, ExFac)
  if (rc /= 0) stop 14
  nBasSh(1) = nBas
  call cho_lk_scf_gpu(rc, 2, LKa, LKb, KLTa, KLTb, nForb, nIorb, MSQa, MSQb, DLT, ExFac, 10, 0.1_8, 1.0_8, &
                      0, nBasSh)
  if (rc /= 0) stop 15
  ! CASSCF: src/fock_util/cho_cas_drv.F90:277 and :285
  call cho_fmcscf_gpu(rc, FI, FA, 0, nI, nA, FactXI, DI, DA, .true., POrb1, nChM, POrb2, POrb3, W, CMO, &
                      nFro, nOrb, 1.0_8)
  if (rc /= 0) stop 16
  call cho_lk_casscf_gpu(rc, DI, DA, FI2, FA2, MSQi, W2, FactXI, nI, nA, nChM, POrb3, Ash2, .true., 10, &
                         0.1_8, 1.0_8, CMO, nFro, nOrb, 1.0_8)
  if (rc /= 0) stop 17
  call jk_gpu_final(rc)

  nbad = 0
  call check(FLTa, nP)
  call check(FLTb, nP)
  call check(LKa, nP)
  call check(LKb, nP)
  call check(KLTa, nP)
  call check(KLTb, nP)
  call check(FI, nP)
  call check(FA, nP)
  call check(W, nW)
  call check(FI2, nP)
  call check(FA2, nP)
  call check(W2, nW)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) FLTa(1:nP), FLTb(1:nP), LKa(1:nP), LKb(1:nP), KLTa(1:nP), KLTb(1:nP)
  write(u) FI(1:nP), FA(1:nP), W(1:nW), FI2(1:nP), FA2(1:nP), W2(1:nW)
  write(u) real(nbad, 8)
  close(u)
  write(*, '(a,i0,a,i0,a,i0)') 'uhf_cas_driver: nBas=', nBas, ' NumCho=', NumCho, ' damaged canaries=', nbad

contains

  subroutine alloc_in(a, m)
    real(8), allocatable, intent(inout) :: a(:)
    integer, intent(in) :: m
    allocate(a(max(m, 1) + nCan))
    a(:) = 0.0_8
    a(max(m, 1)+1:) = canary
  end subroutine alloc_in

  subroutine alloc_out(a, m, init)
    real(8), allocatable, intent(inout) :: a(:)
    integer, intent(in) :: m
    real(8), intent(in), optional :: init(:)
    allocate(a(m + nCan))
    a(:) = 0.0_8
    if (present(init)) a(1:m) = init(1:m)
    a(m+1:) = canary
  end subroutine alloc_out

  subroutine check(a, m)
    real(8), intent(in) :: a(:)
    integer, intent(in) :: m
    nbad = nbad + count(a(m+1:m+nCan) /= canary)
  end subroutine check

end program uhf_cas_driver
