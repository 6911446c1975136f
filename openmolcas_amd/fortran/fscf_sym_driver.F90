!-----------------------------------------------------------------------
! Fortran caller of the symmetry drop-in: two irreps, vectors of JSYM = 1
! and JSYM = 2 in two reduced sets whose maps index the C1 lower triangle.
! Reads from a raw FP64 file
!   nBas(1:2), NumCho(1:2), nRS(1:2), nOcc(1:2), ExFac,
!   map1(nRS(1)), map2(nRS(2)) (0-based C1 LT indices),
!   Lrs1(nRS(1),NumCho(1)), Lrs2(nRS(2),NumCho(2)),
!   Porb (irrep blocks A2(nOcc(s),nBas(s)) back to back), DLT, FLT (irrep LT blocks),
! calls cho_fscf_sym_gpu as ChoSCF_Drv would call CHO_FSCF with nSym = 2 and
! writes FLT.   Usage: fscf_sym_driver <input.bin> <output.bin>
!-----------------------------------------------------------------------
program fscf_sym_driver
  use, intrinsic :: iso_c_binding, only: c_int64_t, c_int32_t
  use molcas_jk_iface
  implicit none
  character(len=512) :: fin, fout
  integer :: u, rc, nBas(2), NumCho(2), nRS(2), nOcc(2), nTri, nPorb, n, i
  real(8) :: hdr(9), ExFac
  real(8), allocatable :: m1(:), m2(:), L1(:), L2(:), Porb(:), DLT(:), FLT(:)
  integer(c_int64_t), allocatable :: map1(:), map2(:)
  integer :: nForb(8, 1), nIorb(8, 1)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  nBas = int(hdr(1:2)); NumCho = int(hdr(3:4)); nRS = int(hdr(5:6)); nOcc = int(hdr(7:8)); ExFac = hdr(9)
  n = nBas(1) + nBas(2)
  nTri = nBas(1)*(nBas(1)+1)/2 + nBas(2)*(nBas(2)+1)/2
  nPorb = nOcc(1)*nBas(1) + nOcc(2)*nBas(2)
  allocate(m1(nRS(1)), m2(nRS(2)), L1(nRS(1)*NumCho(1)), L2(nRS(2)*NumCho(2)), Porb(nPorb), &
           DLT(nTri), FLT(nTri), map1(nRS(1)), map2(nRS(2)))
  read(u) m1
  read(u) m2
  read(u) L1
  read(u) L2
  read(u) Porb
  read(u) DLT
  read(u) FLT
  close(u)
  do i = 1, nRS(1)
    map1(i) = int(m1(i), c_int64_t)
  end do
  do i = 1, nRS(2)
    map2(i) = int(m2(i), c_int64_t)
  end do

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
  nForb = 0
  nIorb = 0
  nIorb(1:2, 1) = nOcc
  call cho_fscf_sym_gpu(rc, 2, nBas, 1, FLT, FLT, nForb, nIorb, Porb, Porb, DLT, ExFac)
  if (rc /= 0) stop 16
  call jk_gpu_final(rc)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) FLT
  close(u)
  write(*, '(a,2i4)') 'fscf_sym_driver: CHO_FSCF nSym=2 nBas=', nBas
end program fscf_sym_driver
