!-----------------------------------------------------------------------
! Fortran caller that takes its vectors from the reference's own Cholesky
! files, as Cho_X_Init + CHO_VECRD would: CHORST and CHRED read by
! mjk_choinfo_read (with the runfile's ISOSHL), the context sized from them
! (nBasT, NumCho summed over irreps), the vectors loaded from CHVEC<iSym>,
! then CHO_FSCF with nSym irreps through cho_fscf_sym_gpu.
! Reads from a raw FP64 file
!   nSym, nBas(1:nSym), ISOSHL(1:nBasT), nOcc(1:nSym), ExFac,
!   Porb (irrep blocks A2(nOcc(s),nBas(s)) back to back), DLT, FLT (irrep LT blocks),
! and writes FLT.   Usage: choinfo_driver <dir with the files> <input.bin> <output.bin>
!-----------------------------------------------------------------------
program choinfo_driver
  use, intrinsic :: iso_c_binding, only: c_ptr, c_int32_t, c_int64_t, c_double, c_char, c_null_char, &
                                         c_null_ptr, c_loc
  use molcas_jk_iface
  implicit none
  character(len=512) :: dir, fin, fout
  character(kind=c_char, len=600), target :: vname(8)
  type(c_ptr) :: info, pv(8)
  integer :: u, rc, nSym, nTri, nPorb, i, s
  integer(c_int32_t) :: nbt, maxred, nshell, adrvec, nsym_f, nbas_f(8)
  integer(c_int64_t) :: numcho, numcho_f(8)
  real(c_double) :: thrcom
  real(8) :: x, ExFac
  integer :: nBas(8), nOcc(8), nForb(8, 1), nIorb(8, 1)
  integer(c_int32_t), allocatable :: isoshl(:)
  real(8), allocatable :: hdr(:), Porb(:), DLT(:), FLT(:)

  call get_command_argument(1, dir)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) x
  nSym = int(x)
  allocate(hdr(nSym))
  read(u) hdr
  nBas = 0
  nBas(1:nSym) = int(hdr)
  allocate(isoshl(sum(nBas)))
  do i = 1, sum(nBas)
    read(u) x
    isoshl(i) = int(x, c_int32_t)
  end do
  read(u) hdr
  nOcc = 0
  nOcc(1:nSym) = int(hdr)
  read(u) ExFac
  nTri = 0
  nPorb = 0
  do s = 1, nSym
    nTri = nTri + nBas(s)*(nBas(s)+1)/2
    nPorb = nPorb + nOcc(s)*nBas(s)
  end do
  allocate(Porb(max(nPorb, 1)), DLT(nTri), FLT(nTri))
  read(u) Porb(1:nPorb)
  read(u) DLT
  read(u) FLT
  close(u)

  rc = int(mjk_choinfo_read(info, trim(dir)//'/CHORST'//c_null_char, trim(dir)//'/CHRED'//c_null_char, isoshl))
  if (rc /= 0) then
    write(*, '(A)') 'choinfo_read: '//jk_error_text(mjk_choinfo_error(info))
    stop 11
  end if
  rc = int(mjk_choinfo_dims(info, nbt, numcho, maxred, nshell, adrvec, thrcom))
  rc = int(mjk_choinfo_symmetry(info, nsym_f, nbas_f, numcho_f))
  if (nsym_f /= nSym .or. any(nbas_f(1:nSym) /= nBas(1:nSym))) stop 12
  call jk_gpu_init(rc, int(nbt), int(numcho))
  if (rc /= 0) stop 13
  do s = 1, nSym
    write(vname(s), '(A,I0,A)') trim(dir)//'/CHVEC', s, c_null_char
    pv(s) = c_loc(vname(s))
  end do
  if (nSym == 1) then
    rc = int(mjk_load_choinfo(jk_ctx, info, vname(1), c_null_ptr, 0_c_int64_t))
  else
    rc = int(mjk_load_choinfo_sym(jk_ctx, info, pv, int(nSym, c_int32_t), c_null_ptr, 0_c_int64_t))
  end if
  if (rc /= 0) then
    write(*, '(A)') 'load_choinfo: '//jk_error_text(mjk_last_error(jk_ctx))
    stop 14
  end if
  rc = int(mjk_choinfo_free(info))

  nForb = 0
  nIorb = 0
  nIorb(1:nSym, 1) = nOcc(1:nSym)
  call cho_fscf_sym_gpu(rc, nSym, nBas(1:nSym), 1, FLT, FLT, nForb, nIorb, Porb, Porb, DLT, ExFac)
  if (rc /= 0) stop 15
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) FLT
  close(u)
  call jk_gpu_final(rc)
  write(*, '(A,I0,A,I0,A)') 'choinfo_driver: nSym=', nSym, ' NumCho=', numcho, ' ok'
end program choinfo_driver
