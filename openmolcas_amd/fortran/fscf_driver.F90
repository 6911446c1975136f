!-----------------------------------------------------------------------
! Fortran caller of the drop-in: reads nBas, NumCho, nOcc, the vectors
! Lrs(nRS,NumCho) (one reduced set covering all pairs), Porb A2(nOcc,nBas),
! DLT and FLT from a raw FP64 file, calls cho_fscf_gpu exactly as
! ChoSCF_Drv calls CHO_FSCF (src/scf/choscf_drv.F90:197) and writes FLT.
! Usage: fscf_driver <input.bin> <output.bin>
!-----------------------------------------------------------------------
program fscf_driver
  use, intrinsic :: iso_c_binding, only: c_int64_t, c_int32_t
  use molcas_jk_iface
  implicit none
  character(len=512) :: fin, fout
  integer :: u, rc, nBas, NumCho, nOcc, nP, i
  real(8) :: hdr(4), ExFac
  real(8), allocatable :: Lrs(:), Porb(:), DLT(:), FLT(:)
  integer(c_int64_t), allocatable :: map(:)
  integer :: nForb(8, 1), nIorb(8, 1)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  nBas = int(hdr(1)); NumCho = int(hdr(2)); nOcc = int(hdr(3)); ExFac = hdr(4)
  nP = nBas*(nBas+1)/2
  allocate(Lrs(nP*NumCho), Porb(nOcc*nBas), DLT(nP), FLT(nP), map(nP))
  read(u) Lrs
  read(u) Porb
  read(u) DLT
  read(u) FLT
  close(u)
  do i = 1, nP
    map(i) = int(i-1, c_int64_t)
  end do

  call jk_gpu_init(rc, nBas, NumCho)
  if (rc /= 0) stop 11
  rc = int(mjk_set_reduced_set(jk_ctx, 0_c_int32_t, int(nP, c_int64_t), map))
  if (rc /= 0) stop 12
  rc = int(mjk_upload_vectors(jk_ctx, 0_c_int32_t, 0_c_int64_t, int(NumCho, c_int64_t), Lrs, 0_c_int32_t))
  if (rc /= 0) stop 13
  nForb = 0
  nIorb = 0
  nIorb(1, 1) = nOcc
  ! nDen = 1: the second-density arguments are not referenced
  call cho_fscf_gpu(rc, 1, FLT, FLT, nForb, nIorb, Porb, Porb, DLT, ExFac)
  if (rc /= 0) stop 14
  call jk_gpu_final(rc)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) FLT
  close(u)
  write(*, '(a,i0,a,i0,a,i0)') 'fscf_driver: CHO_FSCF nBas=', nBas, ' NumCho=', NumCho, ' nOcc=', nOcc
end program fscf_driver
