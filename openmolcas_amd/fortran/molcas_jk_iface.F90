!-----------------------------------------------------------------------
! ISO_C_BINDING interface to libmolcas_jk (include/molcas_jk.h) and
! Fortran drop-ins with the argument lists of OpenMolcas's
!   CHO_FSCF   (src/fock_util/cho_fscf.F90:14)
!   CHO_FMCSCF (src/fock_util/cho_fmcscf.F90:14)
! for nSym = 1.  The reference passes arrays of DSBA_Type, one separately
! allocated object per density (FLT(1:nDen), Porb(1:nDen), KLT(1:nDen),
! src/scf/choscf_drv.F90:285-291,401-402; FI/FA and DI/DA,
! src/fock_util/cho_cas_drv.F90:42-43,101-102).  The shims take the %A0
! storage of each density as a dummy argument of its own (FLT1, FLT2, ...)
! and hand the library one pointer per density; nothing is assumed about
! where the arrays of two densities sit.  For nDen = 1 the second-density
! dummies are not referenced (pass the first array again).  rc keeps the
! reference's meaning (0 ok, 33 memory, 77 short read); the GPU context
! plays the role of the vector buffer of Cho_X_Init.  Every shim takes an
! optional precision (MJK_PREC_*; default FP64).
!-----------------------------------------------------------------------
module molcas_jk_iface
  use, intrinsic :: iso_c_binding, only: c_ptr, c_int, c_int32_t, c_int64_t, c_double, c_null_ptr, &
                                         c_char, c_loc
  implicit none
  private
  public :: mjk_create, mjk_destroy, mjk_set_reduced_set, mjk_upload_vectors, mjk_upload_vectors_file, mjk_cho_fscf, &
            mjk_cho_fmcscf, mjk_cho_lk_scf, mjk_cho_fock_dft_red, cho_fscf_gpu, cho_lk_scf_gpu, &
            mjk_cho_fscf_sym, cho_fscf_sym_gpu, mjk_cho_lk_scf_sym, cho_lk_scf_sym_gpu, &
            cho_fock_dft_red_gpu, cho_fmcscf_gpu, mjk_cho_fmcscf_sym, cho_fmcscf_sym_gpu, mjk_cho_lk_casscf, &
            cho_lk_casscf_gpu, mjk_cho_lk_casscf_sym, cho_lk_casscf_sym_gpu, jk_gpu_init, jk_gpu_final, &
            mjk_set_ladder, mjk_last_error, mjk_set_lk_screening, mjk_set_shells, mjk_set_lk_diag, &
            mjk_choinfo_read, mjk_choinfo_error, mjk_choinfo_free, mjk_choinfo_dims, mjk_choinfo_symmetry, &
            mjk_load_choinfo, mjk_load_choinfo_sym, jk_error_text

  integer(c_int32_t), parameter, public :: MJK_PREC_FP64 = 0, MJK_PREC_FP32 = 1, MJK_PREC_BF16X3 = 2, &
                                           MJK_PREC_LADDER = 3
  type(c_ptr), save, public :: jk_ctx = c_null_ptr

  interface
    integer(c_int) function mjk_create(ctx, nbas, nchol) bind(C, name='mjk_create')
      import :: c_ptr, c_int, c_int32_t, c_int64_t
      type(c_ptr), intent(out) :: ctx
      integer(c_int32_t), value :: nbas
      integer(c_int64_t), value :: nchol
    end function mjk_create
    integer(c_int) function mjk_destroy(ctx) bind(C, name='mjk_destroy')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function mjk_destroy
    integer(c_int) function mjk_set_reduced_set(ctx, jred, nrs, rs_to_lt) bind(C, name='mjk_set_reduced_set')
      import :: c_ptr, c_int, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: jred
      integer(c_int64_t), value :: nrs
      integer(c_int64_t), intent(in) :: rs_to_lt(*)
    end function mjk_set_reduced_set
    integer(c_int) function mjk_upload_vectors(ctx, jred, jvec, jnum, lrs, on_device) &
        bind(C, name='mjk_upload_vectors')
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: jred, on_device
      integer(c_int64_t), value :: jvec, jnum
      real(c_double), intent(in) :: lrs(*)
    end function mjk_upload_vectors
    ! path: file name terminated by c_null_char; iadr(i) = InfVec(jvec+i,3,iSym);
    ! mbl = MBL(Lu): 8 (CHO_ADRVEC=1, word addresses) or 512 (CHO_ADRVEC=2);
    ! split_bytes = 0, or the piece size of a unit split by MOLCAS_MAXFILESIZE
    integer(c_int) function mjk_upload_vectors_file(ctx, path, jred, jvec, jnum, iadr, mbl, split_bytes) &
        bind(C, name='mjk_upload_vectors_file')
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: jred, mbl
      integer(c_int64_t), value :: jvec, jnum, split_bytes
      integer(c_int64_t), intent(in) :: iadr(*)
    end function mjk_upload_vectors_file
    ! text of the last error (a C string: jk_error_text turns it into a Fortran one)
    type(c_ptr) function mjk_last_error(ctx) bind(C, name='mjk_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function mjk_last_error
    ! LK state: screening on/off with ThrCom, the shells nBasSh(1:nShell) and the
    ! integral diagonal CHO_IODIAG reads (LT order)
    integer(c_int) function mjk_set_lk_screening(ctx, enable, thrcom) bind(C, name='mjk_set_lk_screening')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
      real(c_double), value :: thrcom
    end function mjk_set_lk_screening
    integer(c_int) function mjk_set_shells(ctx, nshell, nbassh) bind(C, name='mjk_set_shells')
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nshell
      integer(c_int32_t), intent(in) :: nbassh(*)
    end function mjk_set_shells
    integer(c_int) function mjk_set_lk_diag(ctx, diag, maxred, estimate, update) bind(C, name='mjk_set_lk_diag')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: diag(*)
      integer(c_int32_t), value :: maxred, estimate, update
    end function mjk_set_lk_diag
    ! the reference's Cholesky files (CHORST, CHRED, CHVEC<iSym>, CHODIAG); names
    ! end in c_null_char, isoshl = the runfile array ISOSHL (nBasT values)
    integer(c_int) function mjk_choinfo_read(info, chorst, chred, isoshl) bind(C, name='mjk_choinfo_read')
      import :: c_ptr, c_int, c_int32_t, c_char
      type(c_ptr), intent(out) :: info
      character(kind=c_char), intent(in) :: chorst(*), chred(*)
      integer(c_int32_t), intent(in) :: isoshl(*)
    end function mjk_choinfo_read
    type(c_ptr) function mjk_choinfo_error(info) bind(C, name='mjk_choinfo_error')
      import :: c_ptr
      type(c_ptr), value :: info
    end function mjk_choinfo_error
    integer(c_int) function mjk_choinfo_free(info) bind(C, name='mjk_choinfo_free')
      import :: c_ptr, c_int
      type(c_ptr), value :: info
    end function mjk_choinfo_free
    integer(c_int) function mjk_choinfo_dims(info, nbas, numcho, maxred, nshell, adrvec, thrcom) &
        bind(C, name='mjk_choinfo_dims')
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: info
      integer(c_int32_t), intent(out) :: nbas, maxred, nshell, adrvec
      integer(c_int64_t), intent(out) :: numcho
      real(c_double), intent(out) :: thrcom
    end function mjk_choinfo_dims
    integer(c_int) function mjk_choinfo_symmetry(info, nsym, nbas_sym, numcho_sym) &
        bind(C, name='mjk_choinfo_symmetry')
      import :: c_ptr, c_int, c_int32_t, c_int64_t
      type(c_ptr), value :: info
      integer(c_int32_t), intent(out) :: nsym, nbas_sym(8)
      integer(c_int64_t), intent(out) :: numcho_sym(8)
    end function mjk_choinfo_symmetry
    ! chodiag: c_loc of a c_null_char-terminated name, or c_null_ptr
    integer(c_int) function mjk_load_choinfo(ctx, info, chvec, chodiag, split_bytes) bind(C, name='mjk_load_choinfo')
      import :: c_ptr, c_int, c_int64_t, c_char
      type(c_ptr), value :: ctx, info, chodiag
      character(kind=c_char), intent(in) :: chvec(*)
      integer(c_int64_t), value :: split_bytes
    end function mjk_load_choinfo
    ! chvec(iSym): c_loc of the name of CHVEC<iSym>, nfiles = nSym
    integer(c_int) function mjk_load_choinfo_sym(ctx, info, chvec, nfiles, chodiag, split_bytes) &
        bind(C, name='mjk_load_choinfo_sym')
      import :: c_ptr, c_int, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx, info, chodiag
      type(c_ptr), intent(in) :: chvec(*)
      integer(c_int32_t), value :: nfiles
      integer(c_int64_t), value :: split_bytes
    end function mjk_load_choinfo_sym
    integer(c_int) function mjk_set_ladder(ctx, j_fp32, j_bf16) bind(C, name='mjk_set_ladder')
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: j_fp32, j_bf16
    end function mjk_set_ladder
    ! flt(d), porb(d): c_loc of FLT(d)%A0 and Porb(d)%A0
    integer(c_int) function mjk_cho_fscf(ctx, nden, flt, nforb, niorb, porb, dlt, exfac, prec) &
        bind(C, name='mjk_cho_fscf')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nden, prec
      type(c_ptr), intent(in) :: flt(*), porb(*)
      integer(c_int32_t), intent(in) :: nforb(*), niorb(*)
      real(c_double), intent(in) :: dlt(*)
      real(c_double), value :: exfac
    end function mjk_cho_fscf
    integer(c_int) function mjk_cho_fscf_sym(ctx, nsym, nbas, nden, flt, nforb, niorb, porb, dlt, &
        exfac, prec) bind(C, name='mjk_cho_fscf_sym')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsym, nden, prec
      integer(c_int32_t), intent(in) :: nbas(*), nforb(*), niorb(*)
      type(c_ptr), intent(in) :: flt(*), porb(*)
      real(c_double), intent(in) :: dlt(*)
      real(c_double), value :: exfac
    end function mjk_cho_fscf_sym
    integer(c_int) function mjk_cho_lk_scf_sym(ctx, nsym, nbas, nden, flt, klt, nforb, niorb, porb, plt, &
        factxi, nscreen, dmpk, dfmat, nshell, nbassh, prec) bind(C, name='mjk_cho_lk_scf_sym')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx, klt
      integer(c_int32_t), value :: nsym, nden, nscreen, nshell, prec
      integer(c_int32_t), intent(in) :: nbas(*), nforb(*), niorb(*), nbassh(*)
      type(c_ptr), intent(in) :: flt(*), porb(*)
      real(c_double), intent(in) :: plt(*)
      real(c_double), value :: factxi, dmpk, dfmat
    end function mjk_cho_lk_scf_sym
    integer(c_int) function mjk_cho_lk_casscf(ctx, dlt, flt, msq, w_pwxy, factxi, nfiorb, naorb, nchm, &
        ash1, ash2, do_active, nscreen, dmpk, dfmat, cmo, nfro, norb, exfac, prec) &
        bind(C, name='mjk_cho_lk_casscf')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      type(c_ptr), intent(in) :: dlt(*), flt(*)
      real(c_double), intent(in) :: msq(*), ash1(*), ash2(*), cmo(*)
      real(c_double), intent(inout) :: w_pwxy(*)
      real(c_double), value :: factxi, dmpk, dfmat, exfac
      integer(c_int32_t), value :: nfiorb, naorb, nchm, do_active, nscreen, nfro, norb, prec
    end function mjk_cho_lk_casscf
    integer(c_int) function mjk_cho_fock_dft_red(ctx, dlt, flt) bind(C, name='mjk_cho_fock_dft_red')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: dlt(*)
      real(c_double), intent(inout) :: flt(*)
    end function mjk_cho_fock_dft_red
    integer(c_int) function mjk_cho_lk_scf(ctx, nden, flt, klt, nforb, niorb, porb, plt, factxi, &
        nscreen, dmpk, dfmat, nshell, nbassh, prec) bind(C, name='mjk_cho_lk_scf')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nden, nscreen, nshell, prec
      type(c_ptr), intent(in) :: flt(*), klt(*), porb(*)
      integer(c_int32_t), intent(in) :: nforb(*), niorb(*), nbassh(*)
      real(c_double), intent(in) :: plt(*)
      real(c_double), value :: factxi, dmpk, dfmat
    end function mjk_cho_lk_scf
    integer(c_int) function mjk_cho_fmcscf(ctx, flt, nforb, niorb, naorb, factxi, dlt, do_active, &
        porb1, nchm, porb2, porb3, w_pwxy, cmo, nfro, norb, exfac, prec) bind(C, name='mjk_cho_fmcscf')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      type(c_ptr), intent(in) :: flt(*), dlt(*)
      integer(c_int32_t), value :: nforb, niorb, naorb, do_active, nchm, nfro, norb, prec
      real(c_double), value :: factxi, exfac
      real(c_double), intent(in) :: porb1(*), porb2(*), porb3(*), cmo(*)
      real(c_double), intent(inout) :: w_pwxy(*)
    end function mjk_cho_fmcscf
    integer(c_int) function mjk_cho_fmcscf_sym(ctx, nsym, nbas, flt, nforb, niorb, naorb, factxi, dlt, &
        do_active, porb1, nchm, porb2, porb3, w_pwxy, cmo, nfro, norb, exfac, prec) &
        bind(C, name='mjk_cho_fmcscf_sym')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsym, do_active, prec
      integer(c_int32_t), intent(in) :: nbas(*), nforb(*), niorb(*), naorb(*), nchm(*), nfro(*), norb(*)
      type(c_ptr), intent(in) :: flt(*), dlt(*)
      real(c_double), value :: factxi, exfac
      real(c_double), intent(in) :: porb1(*), porb2(*), porb3(*), cmo(*)
      real(c_double), intent(inout) :: w_pwxy(*)
    end function mjk_cho_fmcscf_sym
    integer(c_int) function mjk_cho_lk_casscf_sym(ctx, nsym, nbas, dlt, flt, msq, w_pwxy, factxi, nchi, naorb, &
        nchm, ash1, ash2, do_active, nscreen, dmpk, dfmat, cmo, nfro, norb, exfac, prec) &
        bind(C, name='mjk_cho_lk_casscf_sym')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsym, do_active, nscreen, prec
      integer(c_int32_t), intent(in) :: nbas(*), nchi(*), naorb(*), nchm(*), nfro(*), norb(*)
      type(c_ptr), intent(in) :: dlt(*), flt(*)
      real(c_double), intent(in) :: msq(*), ash1(*), ash2(*), cmo(*)
      real(c_double), intent(inout) :: w_pwxy(*)
      real(c_double), value :: factxi, dmpk, dfmat, exfac
    end function mjk_cho_lk_casscf_sym
  end interface

contains

  ! Cho_X_Init counterpart: one context for nBas functions and NumCho vectors.
  subroutine jk_gpu_init(rc, nBas, NumCho)
    integer, intent(out) :: rc
    integer, intent(in) :: nBas, NumCho
    rc = int(mjk_create(jk_ctx, int(nBas, c_int32_t), int(NumCho, c_int64_t)))
  end subroutine jk_gpu_init

  ! Cho_X_Final counterpart.
  subroutine jk_gpu_final(rc)
    integer, intent(out) :: rc
    rc = int(mjk_destroy(jk_ctx))
    jk_ctx = c_null_ptr
  end subroutine jk_gpu_final

  ! A C string (mjk_last_error, mjk_choinfo_error) as a Fortran string.
  function jk_error_text(p) result(msg)
    use, intrinsic :: iso_c_binding, only: c_f_pointer, c_associated, c_null_char
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: ch(:)
    integer :: k
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, ch, [4096])
    k = 0
    do while (k < 4096)
      if (ch(k + 1) == c_null_char) exit
      k = k + 1
    end do
    allocate(character(len=k) :: msg)
    do k = 1, len(msg)
      msg(k:k) = ch(k)
    end do
  end function jk_error_text

  integer(c_int32_t) function prec_of(prec)
    integer, intent(in), optional :: prec
    prec_of = MJK_PREC_FP64
    if (present(prec)) prec_of = int(prec, c_int32_t)
  end function prec_of

  ! CHO_FSCF(rc,nDen,FLT,nForb,nIorb,Porb,DLT,ExFac), nSym = 1, called as
  !   call cho_fscf_gpu(rc,nDen,FLT(1)%A0,FLT(2)%A0,nForb,nIorb,Porb(1)%A0,Porb(2)%A0,DLT(1)%A0,ExFac)
  ! FLTd += J[DLT] - ExFac * sum_Jk X X^T with Porbd = A2(nOcc,nBas); the
  ! second-density arrays are referenced only when nDen = 2.
  subroutine cho_fscf_gpu(rc, nDen, FLT1, FLT2, nForb, nIorb, Porb1, Porb2, DLT, ExFac, prec)
    integer, intent(out) :: rc
    integer, intent(in) :: nDen, nForb(8, *), nIorb(8, *)
    real(c_double), intent(inout), target :: FLT1(*), FLT2(*)
    real(c_double), intent(in), target :: Porb1(*), Porb2(*)
    real(c_double), intent(in) :: DLT(*)
    real(8), intent(in) :: ExFac
    integer, intent(in), optional :: prec
    integer(c_int32_t) :: nf(2), ni(2)
    type(c_ptr) :: pf(2), pp(2)
    integer :: d
    nf = 0
    ni = 0
    do d = 1, min(nDen, 2)
      nf(d) = int(nForb(1, d), c_int32_t)
      ni(d) = int(nIorb(1, d), c_int32_t)
    end do
    pf(1) = c_loc(FLT1(1))
    pp(1) = c_loc(Porb1(1))
    pf(2) = pf(1)
    pp(2) = pp(1)
    if (nDen == 2) then
      pf(2) = c_loc(FLT2(1))
      pp(2) = c_loc(Porb2(1))
    end if
    rc = int(mjk_cho_fscf(jk_ctx, int(nDen, c_int32_t), pf, nf, ni, pp, DLT, real(ExFac, c_double), &
                          prec_of(prec)))
  end subroutine cho_fscf_gpu

  ! CHO_FSCF with nSym irreps: FLTd, DLT and Porbd are the %A0 memory of the
  ! DSBA types (irrep blocks back to back), nForb and nIorb the reference's
  ! (8,nDen) arrays.
  subroutine cho_fscf_sym_gpu(rc, nSym, nBas, nDen, FLT1, FLT2, nForb, nIorb, Porb1, Porb2, DLT, ExFac, prec)
    integer, intent(out) :: rc
    integer, intent(in) :: nSym, nBas(nSym), nDen, nForb(8, *), nIorb(8, *)
    real(c_double), intent(inout), target :: FLT1(*), FLT2(*)
    real(c_double), intent(in), target :: Porb1(*), Porb2(*)
    real(c_double), intent(in) :: DLT(*)
    real(8), intent(in) :: ExFac
    integer, intent(in), optional :: prec
    integer(c_int32_t) :: nb(8), nf(8, 2), ni(8, 2)
    type(c_ptr) :: pf(2), pp(2)
    nb = 0
    nf = 0
    ni = 0
    nb(1:nSym) = int(nBas(1:nSym), c_int32_t)
    nf(:, 1:min(nDen, 2)) = int(nForb(:, 1:min(nDen, 2)), c_int32_t)
    ni(:, 1:min(nDen, 2)) = int(nIorb(:, 1:min(nDen, 2)), c_int32_t)
    pf(1) = c_loc(FLT1(1))
    pp(1) = c_loc(Porb1(1))
    pf(2) = pf(1)
    pp(2) = pp(1)
    if (nDen == 2) then
      pf(2) = c_loc(FLT2(1))
      pp(2) = c_loc(Porb2(1))
    end if
    rc = int(mjk_cho_fscf_sym(jk_ctx, int(nSym, c_int32_t), nb, int(nDen, c_int32_t), pf, nf, ni, &
                              pp, DLT, real(ExFac, c_double), prec_of(prec)))
  end subroutine cho_fscf_sym_gpu

  ! CHO_LK_SCF with nSym irreps: FLTd, PLT and Porbd (MSQ A2(nBas(s),nBas(s))
  ! blocks) as the %A0 memory of the DSBA types; KLT is not returned.
  subroutine cho_lk_scf_sym_gpu(rc, nSym, nBas, nDen, FLT1, FLT2, nForb, nIorb, Porb1, Porb2, PLT, FactXI, &
                                nScreen, dmpk, dFmat, prec, KLT1, KLT2, nShell, nBasSh)
    integer, intent(out) :: rc
    integer, intent(in) :: nSym, nBas(nSym), nDen, nForb(8, *), nIorb(8, *), nScreen
    real(c_double), intent(inout), target :: FLT1(*), FLT2(*)
    real(c_double), intent(in), target :: Porb1(*), Porb2(*)
    real(c_double), intent(in) :: PLT(*)
    real(8), intent(in) :: FactXI, dmpk, dFmat
    integer, intent(in), optional :: prec
    ! KLT(1:nDen)%A0 (irrep blocks in shell-pair storage) and nBasSh(nSym, nShell)
    real(c_double), intent(inout), target, optional :: KLT1(*), KLT2(*)
    integer, intent(in), optional :: nShell, nBasSh(nSym, *)
    integer(c_int32_t) :: nb(8), nf(8, 2), ni(8, 2), nsh
    integer(c_int32_t), allocatable :: sh(:)
    type(c_ptr) :: pf(2), pp(2)
    type(c_ptr), target :: pk(2)
    type(c_ptr) :: kp
    nb = 0
    nf = 0
    ni = 0
    nb(1:nSym) = int(nBas(1:nSym), c_int32_t)
    nf(:, 1:min(nDen, 2)) = int(nForb(:, 1:min(nDen, 2)), c_int32_t)
    ni(:, 1:min(nDen, 2)) = int(nIorb(:, 1:min(nDen, 2)), c_int32_t)
    pf(1) = c_loc(FLT1(1))
    pp(1) = c_loc(Porb1(1))
    pf(2) = pf(1)
    pp(2) = pp(1)
    if (nDen == 2) then
      pf(2) = c_loc(FLT2(1))
      pp(2) = c_loc(Porb2(1))
    end if
    kp = c_null_ptr
    if (present(KLT1)) then
      pk(1) = c_loc(KLT1(1))
      pk(2) = pk(1)
      if (nDen == 2 .and. present(KLT2)) pk(2) = c_loc(KLT2(1))
      kp = c_loc(pk)
    end if
    nsh = 0
    if (present(nShell) .and. present(nBasSh)) nsh = int(nShell, c_int32_t)
    allocate(sh(max(1, nSym*nsh)))
    sh = 0
    if (nsh > 0) sh(1:nSym*nsh) = int(reshape(nBasSh(1:nSym, 1:nsh), [nSym*nsh]), c_int32_t)
    rc = int(mjk_cho_lk_scf_sym(jk_ctx, int(nSym, c_int32_t), nb, int(nDen, c_int32_t), pf, kp, nf, ni, &
                                pp, PLT, real(FactXI, c_double), int(nScreen, c_int32_t), &
                                real(dmpk, c_double), real(dFmat, c_double), nsh, sh, prec_of(prec)))
    deallocate(sh)
  end subroutine cho_lk_scf_sym_gpu

  ! CHO_FOCK_DFT_RED(irc,DLT,FLT), nSym = 1: FLT = J[DLT] (pure DFT, no exchange).
  subroutine cho_fock_dft_red_gpu(irc, DLT, FLT)
    integer, intent(out) :: irc
    real(8), intent(in) :: DLT(*)
    real(8), intent(inout) :: FLT(*)
    irc = int(mjk_cho_fock_dft_red(jk_ctx, DLT, FLT))
  end subroutine cho_fock_dft_red_gpu

  ! CHO_LK_SCF(rc,nDen,FLT,KLT,nForb,nIorb,Porb,PLT,FactXI,nScreen,dmpk,dFmat), nSym = 1:
  ! FLTd += J[PLT] - FactXI * K[Porbd], Porbd = MSQ A2(nBas,nBas); KLTd accumulates
  ! -FactXI*K in shell-pair blocks of the nShell shells nBasSh.
  subroutine cho_lk_scf_gpu(rc, nDen, FLT1, FLT2, KLT1, KLT2, nForb, nIorb, Porb1, Porb2, PLT, FactXI, &
                            nScreen, dmpk, dFmat, nShell, nBasSh, prec)
    integer, intent(out) :: rc
    integer, intent(in) :: nDen, nForb(8, *), nIorb(8, *), nScreen, nShell, nBasSh(*)
    real(c_double), intent(inout), target :: FLT1(*), FLT2(*), KLT1(*), KLT2(*)
    real(c_double), intent(in), target :: Porb1(*), Porb2(*)
    real(c_double), intent(in) :: PLT(*)
    real(8), intent(in) :: FactXI, dmpk, dFmat
    integer, intent(in), optional :: prec
    integer(c_int32_t) :: nf(2), ni(2)
    integer(c_int32_t), allocatable :: sh(:)
    type(c_ptr) :: pf(2), pk(2), pp(2)
    integer :: d
    nf = 0
    ni = 0
    do d = 1, min(nDen, 2)
      nf(d) = int(nForb(1, d), c_int32_t)
      ni(d) = int(nIorb(1, d), c_int32_t)
    end do
    pf(1) = c_loc(FLT1(1))
    pk(1) = c_loc(KLT1(1))
    pp(1) = c_loc(Porb1(1))
    pf(2) = pf(1)
    pk(2) = pk(1)
    pp(2) = pp(1)
    if (nDen == 2) then
      pf(2) = c_loc(FLT2(1))
      pk(2) = c_loc(KLT2(1))
      pp(2) = c_loc(Porb2(1))
    end if
    allocate(sh(max(nShell, 1)))
    sh = 0
    if (nShell > 0) sh(1:nShell) = int(nBasSh(1:nShell), c_int32_t)
    rc = int(mjk_cho_lk_scf(jk_ctx, int(nDen, c_int32_t), pf, pk, nf, ni, pp, PLT, real(FactXI, c_double), &
                            int(nScreen, c_int32_t), real(dmpk, c_double), real(dFmat, c_double), &
                            int(nShell, c_int32_t), sh, prec_of(prec)))
    deallocate(sh)
  end subroutine cho_lk_scf_gpu

  ! CHO_LK_CASSCF(DLT,FLT,MSQ,W_PWXY,FactXI,nFIorb,nAorb,nChM,Ash,DoActive,nScreen,dmpk,dFmat,CMO,ExFac),
  ! nSym = 1, called with DLT(1)%A0, DLT(2)%A0, FLT(1)%A0, FLT(2)%A0 (DI, DA, FI, FA);
  ! Ash1 = active MOs A2(nAorb,nBas), Ash2 = DA vectors A2(nBas,nBas).
  subroutine cho_lk_casscf_gpu(rc, DLT1, DLT2, FLT1, FLT2, MSQ, W_PWXY, FactXI, nFIorb, nAorb, nChM, Ash1, &
                               Ash2, DoActive, nScreen, dmpk, dFmat, CMO, nFro, nOrb, ExFac, prec)
    integer, intent(out) :: rc
    integer, intent(in) :: nFIorb, nAorb, nChM, nScreen, nFro, nOrb
    logical, intent(in) :: DoActive
    real(c_double), intent(in), target :: DLT1(*), DLT2(*)
    real(c_double), intent(inout), target :: FLT1(*), FLT2(*)
    real(8), intent(in) :: MSQ(*), Ash1(*), Ash2(*), CMO(*), FactXI, dmpk, dFmat, ExFac
    real(8), intent(inout) :: W_PWXY(*)
    integer, intent(in), optional :: prec
    integer(c_int32_t) :: act
    type(c_ptr) :: pd(2), pf(2)
    act = 0
    pd(1) = c_loc(DLT1(1))
    pf(1) = c_loc(FLT1(1))
    pd(2) = pd(1)
    pf(2) = pf(1)
    if (DoActive) then
      act = 1
      pd(2) = c_loc(DLT2(1))
      pf(2) = c_loc(FLT2(1))
    end if
    rc = int(mjk_cho_lk_casscf(jk_ctx, pd, pf, MSQ, W_PWXY, real(FactXI, c_double), int(nFIorb, c_int32_t), &
                               int(nAorb, c_int32_t), int(nChM, c_int32_t), Ash1, Ash2, act, int(nScreen, c_int32_t), &
                               real(dmpk, c_double), real(dFmat, c_double), CMO, int(nFro, c_int32_t), &
                               int(nOrb, c_int32_t), real(ExFac, c_double), prec_of(prec)))
  end subroutine cho_lk_casscf_gpu

  ! CHO_FMCSCF(rc,FLT,nForb,nIorb,nAorb,FactXI,DLT,DoActive,POrb,nChM,W_PWXY,CMO,ExFac), nSym = 1,
  ! called with FLT(1)%A0, FLT(2)%A0 (FI, FA) and DLT(1)%A0, DLT(2)%A0 (DI, DA).
  subroutine cho_fmcscf_gpu(rc, FLT1, FLT2, nForb, nIorb, nAorb, FactXI, DLT1, DLT2, DoActive, POrb1, nChM, &
                            POrb2, POrb3, W_PWXY, CMO, nFro, nOrb, ExFac, prec)
    integer, intent(out) :: rc
    integer, intent(in) :: nForb, nIorb, nAorb, nChM, nFro, nOrb
    logical, intent(in) :: DoActive
    real(c_double), intent(inout), target :: FLT1(*), FLT2(*)
    real(c_double), intent(in), target :: DLT1(*), DLT2(*)
    real(8), intent(inout) :: W_PWXY(*)
    real(8), intent(in) :: FactXI, POrb1(*), POrb2(*), POrb3(*), CMO(*), ExFac
    integer, intent(in), optional :: prec
    integer(c_int32_t) :: act
    type(c_ptr) :: pd(2), pf(2)
    act = 0
    pd(1) = c_loc(DLT1(1))
    pf(1) = c_loc(FLT1(1))
    pd(2) = pd(1)
    pf(2) = pf(1)
    if (DoActive) then
      act = 1
      pd(2) = c_loc(DLT2(1))
      pf(2) = c_loc(FLT2(1))
    end if
    rc = int(mjk_cho_fmcscf(jk_ctx, pf, int(nForb, c_int32_t), int(nIorb, c_int32_t), int(nAorb, c_int32_t), &
                            real(FactXI, c_double), pd, act, POrb1, int(nChM, c_int32_t), POrb2, POrb3, W_PWXY, &
                            CMO, int(nFro, c_int32_t), int(nOrb, c_int32_t), real(ExFac, c_double), prec_of(prec)))
  end subroutine cho_fmcscf_gpu

  ! CHO_FMCSCF with nSym irreps (src/fock_util/cho_cas_drv.F90:277): FLT1/FLT2,
  ! DLT1/DLT2 are FLT(1:2)%A0, DLT(1:2)%A0 (irrep LT blocks back to back),
  ! POrb1..3 the POrb(1:3)%A0 and CMO the CMO%A0 memory; the counts are the
  ! reference's 8-element arrays.  W_PWXY in the CHO_eval_waxy block order.
  subroutine cho_fmcscf_sym_gpu(rc, nSym, nBas, FLT1, FLT2, nForb, nIorb, nAorb, FactXI, DLT1, DLT2, DoActive, &
                                POrb1, nChM, POrb2, POrb3, W_PWXY, CMO, nFro, nOrb, ExFac, prec)
    integer, intent(out) :: rc
    integer, intent(in) :: nSym, nBas(nSym), nForb(8), nIorb(8), nAorb(8), nChM(8), nFro(8), nOrb(8)
    logical, intent(in) :: DoActive
    real(c_double), intent(inout), target :: FLT1(*), FLT2(*)
    real(c_double), intent(in), target :: DLT1(*), DLT2(*)
    real(8), intent(inout) :: W_PWXY(*)
    real(8), intent(in) :: FactXI, POrb1(*), POrb2(*), POrb3(*), CMO(*), ExFac
    integer, intent(in), optional :: prec
    integer(c_int32_t) :: act, nb(8), nf(8), ni(8), na(8), nc(8), nr(8), no(8)
    type(c_ptr) :: pd(2), pf(2)
    nb = 0
    nb(1:nSym) = int(nBas(1:nSym), c_int32_t)
    nf = int(nForb, c_int32_t)
    ni = int(nIorb, c_int32_t)
    na = int(nAorb, c_int32_t)
    nc = int(nChM, c_int32_t)
    nr = int(nFro, c_int32_t)
    no = int(nOrb, c_int32_t)
    act = 0
    pd(1) = c_loc(DLT1(1))
    pf(1) = c_loc(FLT1(1))
    pd(2) = pd(1)
    pf(2) = pf(1)
    if (DoActive) then
      act = 1
      pd(2) = c_loc(DLT2(1))
      pf(2) = c_loc(FLT2(1))
    end if
    rc = int(mjk_cho_fmcscf_sym(jk_ctx, int(nSym, c_int32_t), nb, pf, nf, ni, na, real(FactXI, c_double), pd, &
                                act, POrb1, nc, POrb2, POrb3, W_PWXY, CMO, nr, no, real(ExFac, c_double), &
                                prec_of(prec)))
  end subroutine cho_fmcscf_sym_gpu

  ! CHO_LK_CASSCF(DLT,FLT,MSQ,W_PWXY,FactXI,nChI,nAorb,nChM,CVa,DoActive,nScreen,dmpK,dFmat,CMO,ExFac)
  ! with nSym irreps (src/fock_util/cho_cas_drv.F90:285), unscreened: DLTd, FLTd
  ! the DLT(d)%A0, FLT(d)%A0 irrep blocks, MSQ / Ash2 the MSQ%A0 and CVa(2)%A0
  ! squares, Ash1 the CVa(1)%A0 blocks A2(nAorb(s),nBas(s)).
  subroutine cho_lk_casscf_sym_gpu(rc, nSym, nBas, DLT1, DLT2, FLT1, FLT2, MSQ, W_PWXY, FactXI, nChI, nAorb, &
                                   nChM, Ash1, Ash2, DoActive, nScreen, dmpk, dFmat, CMO, nFro, nOrb, ExFac, prec)
    integer, intent(out) :: rc
    integer, intent(in) :: nSym, nBas(nSym), nChI(8), nAorb(8), nChM(8), nScreen, nFro(8), nOrb(8)
    logical, intent(in) :: DoActive
    real(c_double), intent(inout), target :: FLT1(*), FLT2(*)
    real(c_double), intent(in), target :: DLT1(*), DLT2(*)
    real(8), intent(inout) :: W_PWXY(*)
    real(8), intent(in) :: MSQ(*), FactXI, Ash1(*), Ash2(*), dmpk, dFmat, CMO(*), ExFac
    integer, intent(in), optional :: prec
    integer(c_int32_t) :: act, nb(8), ni(8), na(8), nc(8), nr(8), no(8)
    type(c_ptr) :: pd(2), pf(2)
    nb = 0
    nb(1:nSym) = int(nBas(1:nSym), c_int32_t)
    ni = int(nChI, c_int32_t)
    na = int(nAorb, c_int32_t)
    nc = int(nChM, c_int32_t)
    nr = int(nFro, c_int32_t)
    no = int(nOrb, c_int32_t)
    act = 0
    pd(1) = c_loc(DLT1(1))
    pf(1) = c_loc(FLT1(1))
    pd(2) = pd(1)
    pf(2) = pf(1)
    if (DoActive) then
      act = 1
      pd(2) = c_loc(DLT2(1))
      pf(2) = c_loc(FLT2(1))
    end if
    rc = int(mjk_cho_lk_casscf_sym(jk_ctx, int(nSym, c_int32_t), nb, pd, pf, MSQ, W_PWXY, real(FactXI, c_double), &
                                   ni, na, nc, Ash1, Ash2, act, int(nScreen, c_int32_t), real(dmpk, c_double), &
                                   real(dFmat, c_double), CMO, nr, no, real(ExFac, c_double), prec_of(prec)))
  end subroutine cho_lk_casscf_sym_gpu

end module molcas_jk_iface
