! Fortran drop-in for RSDFT's Kohn-Sham inner loop on one MI355X.
!
! The generic names keep the reference's module interfaces:
!   hamiltonian(tpsi, htpsi, n, k, s)            hamiltonian_module.f90:44-46,118-122
!   gram_schmidt(n0, n1, k, s)                   gram_schmidt/gram_schmidt_module.f90:35-79
!   subspace_diag(k, s, unk, esp)                subspace_diag/subspace_diag_module.f90:17-19
!   conjugate_gradient(n1, n2, k, s, unk, esp, res)
!     unk(n1:,:,:,:), esp(:,:,:), res(:,:,:)    conjugate_gradient/cg_module.f90:53-61
!     (also unk(:,:), esp(:), res(:) of one k, s slice)
!   esp_calc(wf(:,n_in:), e(n_in:), n_in, k, s)  esp_calc_module.f90:17-68
! and call the C ABI of librsdft_gpu (include/rsdft_gpu.h) through
! ISO_C_BINDING.  Gamma point, one spin: k and s must be 1.
!
! Band storage: the reference keeps unk(ML, MB, k, s) on the host.  By
! default every shim that takes unk copies it to the device store, works
! there and copies the result back, so a caller can swap the modules without
! other changes.  After rsdft_gpu_set_resident(.true.) the device store is
! the master copy and the shims skip the copies (esp and res still come
! back); gram_schmidt, which has no unk argument, works on the attached host
! array (rsdft_gpu_attach_unk) when not resident.
!
! Errors: the library returns a status; a nonzero status prints the
! library's message and stops the program, as stop_program does in the
! reference (ext0/stop_program.f90:1-14).
module rsdft_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private

  public :: hamiltonian, gram_schmidt, subspace_diag, conjugate_gradient, esp_calc
  public :: rsdft_gpu_init, rsdft_gpu_finalize, rsdft_gpu_set_grid, rsdft_gpu_set_vloc
  public :: rsdft_gpu_build_gth, rsdft_gpu_alloc_unk, rsdft_gpu_upload_unk, rsdft_gpu_download_unk
  public :: rsdft_gpu_attach_unk, rsdft_gpu_set_resident, rsdft_gpu_set_cg
  public :: rsdft_gpu_set_local_gth, rsdft_gpu_scf_constants, rsdft_gpu_init_density
  public :: rsdft_gpu_hartree_xc, rsdft_gpu_update_vloc, rsdft_gpu_calc_density
  public :: rsdft_gpu_density_integrals, rsdft_gpu_mix_init, rsdft_gpu_mix_pulay
  public :: rsdft_gpu_normalize_density, rsdft_gpu_get_field, rsdft_gpu_dv
  public :: RSG_FIELD_RHO, RSG_FIELD_RHO_OUT, RSG_FIELD_VION, RSG_FIELD_VH, RSG_FIELD_VXC, RSG_FIELD_VLOC

  integer, parameter :: RSG_FIELD_RHO = 0, RSG_FIELD_RHO_OUT = 1, RSG_FIELD_VION = 2
  integer, parameter :: RSG_FIELD_VH = 3, RSG_FIELD_VXC = 4, RSG_FIELD_VLOC = 5

  type(c_ptr), save :: ctx = c_null_ptr
  logical, save :: resident = .false.
  integer, save :: mb_store = 0
  integer, save :: ncg_ = 4, pc_kind_ = 4, mloop_ = 3
  real(8), pointer, save :: host_unk(:, :) => null()

  interface
    integer(c_int) function rsg_create(h, device) bind(C, name="rsg_create")
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: device
    end function
    integer(c_int) function rsg_destroy(h) bind(C, name="rsg_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    type(c_ptr) function rsg_last_error(h) bind(C, name="rsg_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function rsg_set_grid(h, n1, n2, n3, aa, md) bind(C, name="rsg_set_grid")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: n1, n2, n3, md
      real(c_double), intent(in) :: aa(*)
    end function
    integer(c_int) function rsg_get_grid(h, dv, hgrid, c0, coef) bind(C, name="rsg_get_grid")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: dv, hgrid(*), c0
      type(c_ptr), value :: coef
    end function
    integer(c_int) function rsg_set_vloc(h, v) bind(C, name="rsg_set_vloc")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: v(*)
    end function
    integer(c_int) function rsg_build_gth_projectors(h, natom, frac, kind, nkind, norb, lo, no, rps, hnl) &
        bind(C, name="rsg_build_gth_projectors")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: natom, nkind
      real(c_double), intent(in) :: frac(*), rps(*), hnl(*)
      integer(c_int), intent(in) :: kind(*), norb(*), lo(*), no(*)
    end function
    integer(c_int) function rsg_hpsi_f64_host(h, psi, hpsi, ld, nb) bind(C, name="rsg_hpsi_f64_host")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: psi(*)
      real(c_double), intent(inout) :: hpsi(*)
      integer(c_int64_t), value :: ld
      integer(c_int), value :: nb
    end function
    integer(c_int) function rsg_alloc_bands(h, mb) bind(C, name="rsg_alloc_bands")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: mb
    end function
    integer(c_int) function rsg_upload_bands(h, u, ld, n0, nb) bind(C, name="rsg_upload_bands")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: u(*)
      integer(c_int64_t), value :: ld
      integer(c_int), value :: n0, nb
    end function
    integer(c_int) function rsg_download_bands(h, u, ld, n0, nb) bind(C, name="rsg_download_bands")
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: u(*)
      integer(c_int64_t), value :: ld
      integer(c_int), value :: n0, nb
    end function
    integer(c_int) function rsg_gram_schmidt(h, n0, n1) bind(C, name="rsg_gram_schmidt")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: n0, n1
    end function
    integer(c_int) function rsg_subspace_diag(h, esp) bind(C, name="rsg_subspace_diag")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: esp(*)
    end function
    integer(c_int) function rsg_conjugate_gradient(h, ncg, pc_kind, mloop, esp, res) &
        bind(C, name="rsg_conjugate_gradient")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: ncg, pc_kind, mloop
      real(c_double), intent(inout) :: esp(*), res(*)
    end function
    integer(c_int) function rsg_esp_calc(h, esp) bind(C, name="rsg_esp_calc")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: esp(*)
    end function
    integer(c_int) function rsg_set_local_gth(h, natom, frac, kind, nkind, zps, rloc, cloc) &
        bind(C, name="rsg_set_local_gth")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: natom, nkind
      real(c_double), intent(in) :: frac(*), zps(*), rloc(*), cloc(*)
      integer(c_int), intent(in) :: kind(*)
    end function
    integer(c_int) function rsg_get_scf_constants(h, eewald, nel) bind(C, name="rsg_get_scf_constants")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: eewald, nel
    end function
    integer(c_int) function rsg_get_field(h, field, out) bind(C, name="rsg_get_field")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: field
      real(c_double), intent(inout) :: out(*)
    end function
    integer(c_int) function rsg_init_density(h, width) bind(C, name="rsg_init_density")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: width
    end function
    integer(c_int) function rsg_normalize_density(h) bind(C, name="rsg_normalize_density")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    integer(c_int) function rsg_calc_density(h, occ) bind(C, name="rsg_calc_density")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: occ(*)
    end function
    integer(c_int) function rsg_hartree_xc(h, field, eh, exc) bind(C, name="rsg_hartree_xc")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: field
      real(c_double), intent(out) :: eh, exc
    end function
    integer(c_int) function rsg_update_vloc(h) bind(C, name="rsg_update_vloc")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    integer(c_int) function rsg_density_integrals(h, field, eloc, eion) bind(C, name="rsg_density_integrals")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: field
      real(c_double), intent(out) :: eloc, eion
    end function
    integer(c_int) function rsg_mix_init(h, mmix, beta) bind(C, name="rsg_mix_init")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: mmix
      real(c_double), value :: beta
    end function
    integer(c_int) function rsg_mix_pulay(h) bind(C, name="rsg_mix_pulay")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: c_strlen
    end function
  end interface

  interface hamiltonian
    module procedure d_hamiltonian1
  end interface
  interface gram_schmidt
    module procedure gram_schmidt_gpu
  end interface
  interface subspace_diag
    module procedure subspace_diag_gpu
  end interface
  interface conjugate_gradient
    module procedure conjugate_gradient_gpu, conjugate_gradient_gpu2
  end interface
  interface esp_calc
    module procedure esp_calc_gpu
  end interface

contains

  subroutine check(rc, who)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: who
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = rsg_last_error(ctx)
    if (c_associated(p)) then
      n = int(c_strlen(p))
      call c_f_pointer(p, msg, [n])
      write (*, '(a,": status ",i0,": ",*(a))') who, rc, msg(1:n)
    else
      write (*, '(a,": status ",i0)') who, rc
    end if
    error stop 7
  end subroutine check

  subroutine gamma_only(k, s, who)
    integer, intent(in) :: k, s
    character(*), intent(in) :: who
    if (k /= 1 .or. s /= 1) then
      write (*, '(a,": only k = 1, s = 1 (Gamma point, one spin)")') who
      error stop 7
    end if
  end subroutine gamma_only

  subroutine rsdft_gpu_init(device)
    integer, intent(in), optional :: device
    integer(c_int) :: dev
    dev = 0
    if (present(device)) dev = device
    if (c_associated(ctx)) return
    call check(rsg_create(ctx, dev), "rsdft_gpu_init")
  end subroutine rsdft_gpu_init

  subroutine rsdft_gpu_finalize()
    integer(c_int) :: rc
    if (.not. c_associated(ctx)) return
    rc = rsg_destroy(ctx)
    ctx = c_null_ptr
    mb_store = 0
    resident = .false.
    host_unk => null()
  end subroutine rsdft_gpu_finalize

  ! Lattice aa(3,3) (columns a1, a2, a3, bohr), grid, FD order Md
  ! (rgrid_module / kinetic_module init).
  subroutine rsdft_gpu_set_grid(ngrid, aa, md)
    integer, intent(in) :: ngrid(3), md
    real(8), intent(in) :: aa(3, 3)
    real(c_double) :: a(9)
    a = reshape(aa, [9])
    call check(rsg_set_grid(ctx, ngrid(1), ngrid(2), ngrid(3), a, md), "rsdft_gpu_set_grid")
  end subroutine rsdft_gpu_set_grid

  real(8) function rsdft_gpu_dv()
    real(c_double) :: dv, hg(3), c0
    call check(rsg_get_grid(ctx, dv, hg, c0, c_null_ptr), "rsdft_gpu_dv")
    rsdft_gpu_dv = dv
  end function rsdft_gpu_dv

  subroutine rsdft_gpu_set_vloc(vloc)
    real(8), intent(in) :: vloc(:)
    call check(rsg_set_vloc(ctx, vloc), "rsdft_gpu_set_vloc")
  end subroutine rsdft_gpu_set_vloc

  ! GTH nonlocal tables (prep_ps_nloc2 for GTH).  Per-species arrays are
  ! (4, nkind): orbital iorb of species ik in column ik.
  subroutine rsdft_gpu_build_gth(frac, kind, norb, lo, no, rps, hnl)
    real(8), intent(in) :: frac(:, :), rps(:, :), hnl(:, :)
    integer, intent(in) :: kind(:), norb(:), lo(:, :), no(:, :)
    integer(c_int), allocatable :: k0(:)
    k0 = kind - 1
    call check(rsg_build_gth_projectors(ctx, size(kind), frac, k0, size(norb), norb, lo, no, rps, hnl), &
               "rsdft_gpu_build_gth")
  end subroutine rsdft_gpu_build_gth

  subroutine rsdft_gpu_alloc_unk(mb)
    integer, intent(in) :: mb
    call check(rsg_alloc_bands(ctx, mb), "rsdft_gpu_alloc_unk")
    mb_store = mb
  end subroutine rsdft_gpu_alloc_unk

  subroutine rsdft_gpu_upload_unk(unk)
    real(8), intent(in) :: unk(:, :)
    if (mb_store /= size(unk, 2)) call rsdft_gpu_alloc_unk(size(unk, 2))
    ! the store's band offset n0 is 0-based
    call check(rsg_upload_bands(ctx, unk, int(size(unk, 1), c_int64_t), 0, size(unk, 2)), "rsdft_gpu_upload_unk")
  end subroutine rsdft_gpu_upload_unk

  subroutine rsdft_gpu_download_unk(unk)
    real(8), intent(inout) :: unk(:, :)
    call check(rsg_download_bands(ctx, unk, int(size(unk, 1), c_int64_t), 0, size(unk, 2)), &
               "rsdft_gpu_download_unk")
  end subroutine rsdft_gpu_download_unk

  ! The host array gram_schmidt works on when the bands are not resident.
  subroutine rsdft_gpu_attach_unk(unk)
    real(8), intent(in), target :: unk(:, :)
    host_unk => unk
  end subroutine rsdft_gpu_attach_unk

  subroutine rsdft_gpu_set_resident(flag)
    logical, intent(in) :: flag
    resident = flag
  end subroutine rsdft_gpu_set_resident

  ! NCG, SWPC (4 = hprsdft, 0 = none) and mloop (cg_module.f90:42-50,
  ! cgpc_module.f90:95-96).
  subroutine rsdft_gpu_set_cg(ncg, pc_kind, mloop)
    integer, intent(in) :: ncg, pc_kind, mloop
    ncg_ = ncg
    pc_kind_ = pc_kind
    mloop_ = mloop
  end subroutine rsdft_gpu_set_cg

  subroutine to_device(unk)
    real(8), intent(in) :: unk(:, :)
    if (.not. resident) call rsdft_gpu_upload_unk(unk)
  end subroutine to_device

  subroutine to_host(unk)
    real(8), intent(inout) :: unk(:, :)
    if (.not. resident) call rsdft_gpu_download_unk(unk)
  end subroutine to_host

  ! htpsi = (T + Vloc) tpsi + V_NL tpsi for the nb = size(tpsi,2) bands
  ! starting at band n (hamiltonian_module.f90:118-144); overwrites htpsi.
  subroutine d_hamiltonian1(tpsi, htpsi, n, k, s)
    real(8), intent(in) :: tpsi(:, :)
    real(8), intent(inout) :: htpsi(:, :)
    integer, intent(in) :: n, k, s
    call gamma_only(k, s, "hamiltonian")
    if (n < 1) error stop "hamiltonian: band index n must be >= 1"
    if (any(shape(tpsi) /= shape(htpsi))) error stop "hamiltonian: tpsi and htpsi differ in shape"
    call check(rsg_hpsi_f64_host(ctx, tpsi, htpsi, int(size(tpsi, 1), c_int64_t), size(tpsi, 2)), "hamiltonian")
  end subroutine d_hamiltonian1

  subroutine gram_schmidt_gpu(n0, n1, k, s)
    integer, intent(in) :: n0, n1, k, s
    call gamma_only(k, s, "gram_schmidt")
    if (.not. resident) then
      if (.not. associated(host_unk)) error stop "gram_schmidt: attach unk or make the bands resident"
      call rsdft_gpu_upload_unk(host_unk)
    end if
    call check(rsg_gram_schmidt(ctx, n0, n1), "gram_schmidt")
    if (.not. resident) call rsdft_gpu_download_unk(host_unk)
  end subroutine gram_schmidt_gpu

  subroutine subspace_diag_gpu(k, s, unk, esp)
    integer, intent(in) :: k, s
    real(8), intent(inout) :: unk(:, :)
    real(8), intent(inout) :: esp(:)
    call gamma_only(k, s, "subspace_diag")
    call to_device(unk)
    call check(rsg_subspace_diag(ctx, esp), "subspace_diag")
    call to_host(unk)
  end subroutine subspace_diag_gpu

  ! conjugate_gradient(n1, n2, k, s, unk, esp, res) with the reference's
  ! whole arrays unk(n1:, MB, k, s), esp(MB, k, s), res(MB, k, s)
  ! (cg_module.f90:53-61, called as conjugate_gradient(ML_0,ML_1,k,s,unk,
  ! esp,res) at scf_module.f90:254).  n1:n2 is the rank's grid slice; on one
  ! rank it is the whole grid.
  subroutine conjugate_gradient_gpu(n1, n2, k, s, unk, esp, res)
    integer, intent(in) :: n1, n2, k, s
    real(8), intent(inout) :: unk(n1:, :, :, :)
    real(8), intent(inout) :: esp(:, :, :), res(:, :, :)
    call gamma_only(k, s, "conjugate_gradient")
    if (n2 - n1 + 1 /= size(unk, 1)) error stop "conjugate_gradient: one rank owns the whole grid"
    if (size(esp, 1) /= size(unk, 2) .or. size(res, 1) /= size(unk, 2)) &
      error stop "conjugate_gradient: esp and res need one entry per band"
    call to_device(unk(:, :, k, s))
    call check(rsg_conjugate_gradient(ctx, ncg_, pc_kind_, mloop_, esp(:, k, s), res(:, k, s)), &
               "conjugate_gradient")
    call to_host(unk(:, :, k, s))
  end subroutine conjugate_gradient_gpu

  ! The same on one (k, s) slice: unk(ML, MB), esp(MB), res(MB).
  subroutine conjugate_gradient_gpu2(n1, n2, k, s, unk, esp, res)
    integer, intent(in) :: n1, n2, k, s
    real(8), intent(inout) :: unk(:, :)
    real(8), intent(inout) :: esp(:), res(:)
    call gamma_only(k, s, "conjugate_gradient")
    if (n2 - n1 + 1 /= size(unk, 1)) error stop "conjugate_gradient: one rank owns the whole grid"
    call to_device(unk)
    call check(rsg_conjugate_gradient(ctx, ncg_, pc_kind_, mloop_, esp, res), "conjugate_gradient")
    call to_host(unk)
  end subroutine conjugate_gradient_gpu2

  ! esp_calc(wf, e, n_in, k, s) with wf(:, n_in:) and e(n_in:) the bands
  ! n_in .. n_in+size(e)-1 of the store (d_esp_calc, esp_calc_module.f90:
  ! 17-68, called as esp_calc(unk(:,MB_0:MB_1,k,s), esp(MB_0:MB_1,k,s),
  ! MB_0,k,s) at scf_module.f90:268).  e_n = dV <wf_n|H wf_n>.
  subroutine esp_calc_gpu(wf, e, n_in, k, s)
    integer, intent(in) :: n_in, k, s
    real(8), intent(in) :: wf(:, n_in:)
    real(8), intent(inout) :: e(n_in:)
    real(8), allocatable :: all_e(:)
    integer :: nn
    call gamma_only(k, s, "esp_calc")
    nn = size(e)
    if (size(wf, 2) /= nn) error stop "esp_calc: wf and e hold different band counts"
    if (n_in < 1 .or. n_in + nn - 1 > mb_store) error stop "esp_calc: bands outside the store"
    if (.not. resident) &
      call check(rsg_upload_bands(ctx, wf, int(size(wf, 1), c_int64_t), n_in - 1, nn), "esp_calc")
    allocate (all_e(mb_store))
    call check(rsg_esp_calc(ctx, all_e), "esp_calc")
    e(n_in:n_in + nn - 1) = all_e(n_in:n_in + nn - 1)
  end subroutine esp_calc_gpu

  ! --- the SCF half on the device (density, Hartree, XC, mixing, energies)

  subroutine rsdft_gpu_set_local_gth(frac, kind, zps, rloc, cloc)
    real(8), intent(in) :: frac(:, :), zps(:), rloc(:), cloc(:, :)
    integer, intent(in) :: kind(:)
    integer(c_int), allocatable :: k0(:)
    k0 = kind - 1
    call check(rsg_set_local_gth(ctx, size(kind), frac, k0, size(zps), zps, rloc, cloc), "rsdft_gpu_set_local_gth")
  end subroutine rsdft_gpu_set_local_gth

  subroutine rsdft_gpu_scf_constants(eewald, nelectron)
    real(8), intent(out) :: eewald, nelectron
    call check(rsg_get_scf_constants(ctx, eewald, nelectron), "rsdft_gpu_scf_constants")
  end subroutine rsdft_gpu_scf_constants

  subroutine rsdft_gpu_init_density(width)
    real(8), intent(in) :: width
    call check(rsg_init_density(ctx, width), "rsdft_gpu_init_density")
  end subroutine rsdft_gpu_init_density

  subroutine rsdft_gpu_normalize_density()
    call check(rsg_normalize_density(ctx), "rsdft_gpu_normalize_density")
  end subroutine rsdft_gpu_normalize_density

  subroutine rsdft_gpu_calc_density(occ)
    real(8), intent(in) :: occ(:)
    call check(rsg_calc_density(ctx, occ), "rsdft_gpu_calc_density")
  end subroutine rsdft_gpu_calc_density

  subroutine rsdft_gpu_hartree_xc(field, eh, exc)
    integer, intent(in) :: field
    real(8), intent(out) :: eh, exc
    call check(rsg_hartree_xc(ctx, field, eh, exc), "rsdft_gpu_hartree_xc")
  end subroutine rsdft_gpu_hartree_xc

  subroutine rsdft_gpu_update_vloc()
    call check(rsg_update_vloc(ctx), "rsdft_gpu_update_vloc")
  end subroutine rsdft_gpu_update_vloc

  subroutine rsdft_gpu_density_integrals(field, eloc, eion)
    integer, intent(in) :: field
    real(8), intent(out) :: eloc, eion
    call check(rsg_density_integrals(ctx, field, eloc, eion), "rsdft_gpu_density_integrals")
  end subroutine rsdft_gpu_density_integrals

  subroutine rsdft_gpu_mix_init(mmix, beta)
    integer, intent(in) :: mmix
    real(8), intent(in) :: beta
    call check(rsg_mix_init(ctx, mmix, beta), "rsdft_gpu_mix_init")
  end subroutine rsdft_gpu_mix_init

  subroutine rsdft_gpu_mix_pulay()
    call check(rsg_mix_pulay(ctx), "rsdft_gpu_mix_pulay")
  end subroutine rsdft_gpu_mix_pulay

  subroutine rsdft_gpu_get_field(field, out)
    integer, intent(in) :: field
    real(8), intent(inout) :: out(:)
    call check(rsg_get_field(ctx, field, out), "rsdft_gpu_get_field")
  end subroutine rsdft_gpu_get_field

end module rsdft_gpu
