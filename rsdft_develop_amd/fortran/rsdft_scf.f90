! Kohn-Sham SCF of a diamond Si supercell driven from Fortran through the
! rsdft_gpu module, in the order of RSDFT's calc_scf (scf_module.f90:183-481):
! subspace_diag -> conjugate_gradient -> gram_schmidt -> esp_calc -> Fermi
! occupations -> density -> Hartree + XC -> total energy -> Pulay mixing.
!
! usage: rsdft_scf [nrep [ngrid_per_cell [nband [ekbt [etol [max_iter]]]]]]
!   defaults: 1 36 (2.5 bands per atom) 0.01 1e-10 100
! Prints one line per iteration and, at the end,
!   ETOT <Etot>  FENE <Etot - entropy>  NITER <n>  CONVERGED <T|F>
!   ESP <sorted eigenvalues>
!   HCHECK <max_n |dV <u_n|H u_n> - esp_n|>, H applied by the host-array
!   hamiltonian shim to the downloaded bands
program rsdft_scf
  use rsdft_gpu
  implicit none
  integer, parameter :: md = 6, mmix = 4
  real(8), parameter :: a8 = 10.26d0, beta = 0.3d0, width = 1.0d0
  real(8), parameter :: sites(3, 8) = reshape([0.00d0, 0.00d0, 0.00d0, 0.00d0, 0.50d0, 0.50d0, &
                                               0.50d0, 0.00d0, 0.50d0, 0.50d0, 0.50d0, 0.00d0, &
                                               0.25d0, 0.25d0, 0.25d0, 0.25d0, 0.75d0, 0.75d0, &
                                               0.75d0, 0.25d0, 0.75d0, 0.75d0, 0.75d0, 0.25d0], [3, 8])
  integer :: nrep, npc, nb, max_iter, natom, n, it, i, j, k, a, niter
  real(8) :: ekbt, etol, aa(3, 3), eew, nel, eh_in, exc_in, eh, exc, eloc, eion, eloc_in, eion_in
  real(8) :: eeig, etot, eharris, eprev, de, fene
  real(8), allocatable :: frac(:, :), occ(:), u1(:), u2(:), hunk(:, :)
  ! band storage as wf_module keeps it: unk(ML_0:ML_1, MB, k, s),
  ! esp(MB, k, s), res(MB, k, s); Gamma point, one spin
  real(8), allocatable :: unk(:, :, :, :), esp(:, :, :), res(:, :, :)
  integer :: ML_0, ML_1, MB_0, MB_1, Nband, kk, ss
  real(8) :: dv, hmax
  integer, allocatable :: kind(:), seed(:)
  integer :: norb(1), lo(4, 1), no(4, 1)
  real(8) :: rps(4, 1), hnl(4, 1), zps(1), rloc(1), cloc(4, 1)
  logical :: conv
  character(64) :: arg

  nrep = 1; npc = 36; nb = 0; ekbt = 0.01d0; etol = 1d-10; max_iter = 100
  if (command_argument_count() >= 1) then; call get_command_argument(1, arg); read (arg, *) nrep; end if
  if (command_argument_count() >= 2) then; call get_command_argument(2, arg); read (arg, *) npc; end if
  if (command_argument_count() >= 3) then; call get_command_argument(3, arg); read (arg, *) nb; end if
  if (command_argument_count() >= 4) then; call get_command_argument(4, arg); read (arg, *) ekbt; end if
  if (command_argument_count() >= 5) then; call get_command_argument(5, arg); read (arg, *) etol; end if
  if (command_argument_count() >= 6) then; call get_command_argument(6, arg); read (arg, *) max_iter; end if
  natom = 8 * nrep**3
  if (nb <= 0) nb = (natom * 5) / 2
  n = npc * nrep
  aa = 0d0
  do i = 1, 3
    aa(i, i) = a8 * nrep
  end do
  allocate (frac(3, natom), kind(natom))
  a = 0
  do i = 0, nrep - 1
    do j = 0, nrep - 1
      do k = 0, nrep - 1
        frac(:, a + 1:a + 8) = (sites + spread([dble(i), dble(j), dble(k)], 2, 8)) / nrep
        a = a + 8
      end do
    end do
  end do
  kind = 1
  ! GTH Si: Z = 4, rloc = 0.44, C1 = -7.336103; s (0.422738, 5.906928),
  ! p (0.484278, 2.727013)
  norb = 2
  lo = 0; lo(2, 1) = 1
  no = 0; no(1:2, 1) = 1
  rps = 0d0; rps(1:2, 1) = [0.422738d0, 0.484278d0]
  hnl = 0d0; hnl(1:2, 1) = [5.906928d0, 2.727013d0]
  zps = 4d0; rloc = 0.44d0
  cloc = 0d0; cloc(1, 1) = -7.336103d0

  call rsdft_gpu_init(0)
  call rsdft_gpu_set_grid([n, n, n], aa, md)
  call rsdft_gpu_build_gth(frac, kind, norb, lo, no, rps, hnl)
  call rsdft_gpu_set_local_gth(frac, kind, zps, rloc, cloc)
  call rsdft_gpu_scf_constants(eew, nel)
  call rsdft_gpu_init_density(width)
  call rsdft_gpu_hartree_xc(RSG_FIELD_RHO, eh_in, exc_in)
  call rsdft_gpu_update_vloc()

  ! N(0,1) starting bands (Box-Muller on the intrinsic generator, fixed seed)
  ML_0 = 1; ML_1 = n**3; MB_0 = 1; MB_1 = nb; Nband = nb; kk = 1; ss = 1
  allocate (unk(ML_0:ML_1, nb, 1, 1), esp(nb, 1, 1), res(nb, 1, 1), occ(nb), u1(n**3), u2(n**3))
  call random_seed(size=i)
  allocate (seed(i))
  seed = 20240607
  call random_seed(put=seed)
  do k = 1, nb
    call random_number(u1)
    call random_number(u2)
    unk(:, k, 1, 1) = sqrt(-2d0 * log(1d0 - u1)) * cos(6.283185307179586d0 * u2)
  end do
  call rsdft_gpu_upload_unk(unk(:, :, 1, 1))
  call rsdft_gpu_set_resident(.true.)
  call gram_schmidt(1, Nband, kk, ss)
  call rsdft_gpu_mix_init(mmix, beta)

  conv = .false.
  eprev = 0d0
  niter = 0
  do it = 1, max_iter
    niter = it
    ! the reference's own call lines (scf_module.f90:231, 254, 258, 268)
    call subspace_diag( kk, ss, unk(:,:,kk,ss), esp(:,kk,ss) )
    call conjugate_gradient(ML_0,ML_1,kk,ss,unk,esp,res)
    call gram_schmidt(1,Nband,kk,ss)
    call esp_calc( unk(:,MB_0:MB_1,kk,ss), esp(MB_0:MB_1,kk,ss), MB_0,kk,ss )
    call fermi(esp(:, 1, 1), nel, ekbt, occ)
    eeig = sum(occ * esp(:, 1, 1))
    call rsdft_gpu_density_integrals(RSG_FIELD_RHO, eloc_in, eion_in)
    eharris = eeig - eloc_in + eh_in + exc_in + eion_in + eew
    call rsdft_gpu_calc_density(occ)
    call rsdft_gpu_hartree_xc(RSG_FIELD_RHO_OUT, eh, exc)
    call rsdft_gpu_density_integrals(RSG_FIELD_RHO_OUT, eloc, eion)
    etot = eeig - eloc + eh + exc + eion + eew
    fene = etot - entropy(occ, ekbt)
    de = huge(1d0)
    if (it > 1) de = abs(etot - eprev) / natom
    write (*, '("iter ",i4,"  etot ",f22.12,"  harris ",f22.12,"  de/atom ",es10.3)') it, etot, eharris, de
    if (de < etol) then
      conv = .true.
      exit
    end if
    eprev = etot
    call rsdft_gpu_mix_pulay()
    call rsdft_gpu_normalize_density()
    call rsdft_gpu_hartree_xc(RSG_FIELD_RHO, eh_in, exc_in)
    call rsdft_gpu_update_vloc()
  end do
  call rsdft_gpu_download_unk(unk(:, :, 1, 1))
  allocate (hunk(n**3, nb))
  call hamiltonian(unk(:, :, 1, 1), hunk, 1, 1, 1)
  dv = rsdft_gpu_dv()
  hmax = 0d0
  do k = 1, nb
    hmax = max(hmax, abs(dv * dot_product(unk(:, k, 1, 1), hunk(:, k)) - esp(k, 1, 1)))
  end do
  call sort(esp(:, 1, 1))
  write (*, '("ETOT ",es25.16,"  FENE ",es25.16,"  NITER ",i0,"  CONVERGED ",l1)') etot, fene, niter, conv
  write (*, '("ESP",*(1x,es25.16))') esp(:, 1, 1)
  write (*, '("HCHECK ",es12.4)') hmax
  call rsdft_gpu_finalize()

contains

  ! Spin-unpolarised Fermi-Dirac occupations, chemical potential by bisection
  ! (fermi_dirac_module.f90:11-110).
  subroutine fermi(e, ne, kt, f)
    real(8), intent(in) :: e(:), ne, kt
    real(8), intent(out) :: f(:)
    real(8) :: lo_, hi_, mu
    integer :: m
    lo_ = minval(e) - 1d0
    hi_ = maxval(e) + 1d0
    do m = 1, 300
      mu = 0.5d0 * (lo_ + hi_)
      f = 2d0 / (1d0 + exp(min(max((e - mu) / kt, -700d0), 700d0)))
      if (sum(f) > ne) then
        hi_ = mu
      else
        lo_ = mu
      end if
    end do
    mu = 0.5d0 * (lo_ + hi_)
    f = 2d0 / (1d0 + exp(min(max((e - mu) / kt, -700d0), 700d0)))
  end subroutine fermi

  ! 2 kT sum [-f ln f - (1-f) ln(1-f)], f = occ/2 (fermi_dirac_module.f90:325-350)
  real(8) function entropy(o, kt)
    real(8), intent(in) :: o(:), kt
    real(8) :: f, g
    integer :: m
    entropy = 0d0
    do m = 1, size(o)
      f = 0.5d0 * o(m)
      g = 1d0 - f
      if (f > 0d0) entropy = entropy - f * log(f)
      if (g > 0d0) entropy = entropy - g * log(g)
    end do
    entropy = 2d0 * kt * entropy
  end function entropy

  subroutine sort(x)
    real(8), intent(inout) :: x(:)
    integer :: p, q
    real(8) :: t
    do p = 2, size(x)
      t = x(p)
      q = p - 1
      do while (q >= 1)
        if (x(q) <= t) exit
        x(q + 1) = x(q)
        q = q - 1
      end do
      x(q + 1) = t
    end do
  end subroutine sort

end program rsdft_scf
