!-----------------------------------------------------------------------
! Fortran caller of the CASSCF symmetry drop-in: two irreps, vectors of
! JSYM = 1 and JSYM = 2 in two reduced sets whose maps index the C1 lower
! triangle.  Reads from a raw FP64 file
!   nBas(1:2), NumCho(1:2), nRS(1:2), nForb(1:2), nIorb(1:2), nAorb(1:2),
!   nChM(1:2), nFro(1:2), nOrb(1:2), FactXI, DoActive, nW,
!   map1(nRS(1)), map2(nRS(2)) (0-based C1 LT indices),
!   Lrs1(nRS(1),NumCho(1)), Lrs2(nRS(2),NumCho(2)),
!   DI, DA, FI, FA (irrep LT blocks back to back),
!   POrb(1..3) (irrep blocks A2(nk(s),nBas(s)) back to back),
!   CMO, MSQ, Ash2 (irrep blocks A2(nBas(s),nBas(s)) back to back),
! calls cho_fmcscf_sym_gpu and cho_lk_casscf_sym_gpu as CHO_CAS_DRV would
! call CHO_FMCSCF and CHO_LK_CASSCF with nSym = 2 (FI/FA and DI/DA allocated
! separately; CVa(1) = POrb(3)) and writes FI, FA, W_PWXY of each.
! Usage: fmcscf_sym_driver <input.bin> <output.bin>
!-----------------------------------------------------------------------
program fmcscf_sym_driver
  use, intrinsic :: iso_c_binding, only: c_int64_t, c_int32_t
  use molcas_jk_iface
  implicit none
  character(len=512) :: fin, fout
  integer :: u, rc, nBas(2), NumCho(2), nRS(2), n, i, nTri, nP1, nP2, nP3, nC, nW
  integer :: nForb(8), nIorb(8), nAorb(8), nChM(8), nFro(8), nOrb(8)
  logical :: DoActive
  real(8) :: hdr(21), FactXI
  real(8), allocatable :: m1(:), m2(:), L1(:), L2(:), DI(:), DA(:), FI(:), FA(:)
  real(8), allocatable :: POrb1(:), POrb2(:), POrb3(:), CMO(:), W(:), MSQ(:), Ash2(:), FI2(:), FA2(:), W2(:)
  integer(c_int64_t), allocatable :: map1(:), map2(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  nForb = 0; nIorb = 0; nAorb = 0; nChM = 0; nFro = 0; nOrb = 0
  nBas = int(hdr(1:2)); NumCho = int(hdr(3:4)); nRS = int(hdr(5:6))
  nForb(1:2) = int(hdr(7:8)); nIorb(1:2) = int(hdr(9:10)); nAorb(1:2) = int(hdr(11:12))
  nChM(1:2) = int(hdr(13:14)); nFro(1:2) = int(hdr(15:16)); nOrb(1:2) = int(hdr(17:18))
  FactXI = hdr(19); DoActive = hdr(20) /= 0.0_8; nW = int(hdr(21))
  n = nBas(1) + nBas(2)
  nTri = nBas(1)*(nBas(1)+1)/2 + nBas(2)*(nBas(2)+1)/2
  nP1 = sum((nForb(1:2) + nIorb(1:2))*nBas)
  nP2 = sum(nChM(1:2)*nBas)
  nP3 = sum(nAorb(1:2)*nBas)
  nC = sum(nBas*nBas)
  allocate(m1(nRS(1)), m2(nRS(2)), L1(nRS(1)*NumCho(1)), L2(nRS(2)*NumCho(2)), map1(nRS(1)), map2(nRS(2)))
  allocate(DI(nTri), DA(nTri), FI(nTri), FA(nTri), POrb1(max(1, nP1)), POrb2(max(1, nP2)), &
           POrb3(max(1, nP3)), CMO(nC), W(max(1, nW)), MSQ(nC), Ash2(nC), FI2(nTri), FA2(nTri), &
           W2(max(1, nW)))
  read(u) m1
  read(u) m2
  read(u) L1
  read(u) L2
  read(u) DI
  read(u) DA
  read(u) FI
  read(u) FA
  read(u) POrb1(1:nP1)
  read(u) POrb2(1:nP2)
  read(u) POrb3(1:nP3)
  read(u) CMO
  read(u) MSQ
  read(u) Ash2
  close(u)
  FI2 = FI
  FA2 = FA
  do i = 1, nRS(1)
    map1(i) = int(m1(i), c_int64_t)
  end do
  do i = 1, nRS(2)
    map2(i) = int(m2(i), c_int64_t)
  end do
  W = 0.0_8
  W2 = 0.0_8

  call jk_gpu_init(rc, n, NumCho(1) + NumCho(2))
  if (rc /= 0) stop 11
  rc = int(mjk_set_reduced_set(jk_ctx, 0_c_int32_t, int(nRS(1), c_int64_t), map1))
  if (rc /= 0) stop 12
  rc = int(mjk_set_reduced_set(jk_ctx, 1_c_int32_t, int(nRS(2), c_int64_t), map2))
  if (rc /= 0) stop 13
  rc = int(mjk_upload_vectors(jk_ctx, 0_c_int32_t, 0_c_int64_t, int(NumCho(1), c_int64_t), L1, 0_c_int32_t))
  if (rc /= 0) stop 14
  rc = int(mjk_upload_vectors(jk_ctx, 1_c_int32_t, int(NumCho(1), c_int64_t), int(NumCho(2), c_int64_t), L2, &
                              0_c_int32_t))
  if (rc /= 0) stop 15
  call cho_fmcscf_sym_gpu(rc, 2, nBas, FI, FA, nForb, nIorb, nAorb, FactXI, DI, DA, DoActive, POrb1, nChM, &
                          POrb2, POrb3, W, CMO, nFro, nOrb, 1.0_8)
  if (rc /= 0) stop 16
  call cho_lk_casscf_sym_gpu(rc, 2, nBas, DI, DA, FI2, FA2, MSQ, W2, FactXI, nForb + nIorb, nAorb, nChM, POrb3, &
                             Ash2, DoActive, 10, 0.1_8, 1.0_8, CMO, nFro, nOrb, 1.0_8)
  if (rc /= 0) stop 17
  call jk_gpu_final(rc)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) FI, FA, W(1:nW), FI2, FA2, W2(1:nW)
  close(u)
  write(*, '(a,2i4)') 'fmcscf_sym_driver: CHO_FMCSCF, CHO_LK_CASSCF nSym=2 nBas=', nBas
end program fmcscf_sym_driver
