!> ref_shim -- TEST INFRASTRUCTURE ONLY.  C-callable entry points into the
!> reference's own Fortran routines, compiled from the reference's sources
!> (oracle/ref.mk) into oracle/_ref/libbrawl_ref.so.  Only tests/ load it,
!> to check the C restatement (liboracle.so) against the reference itself on
!> inputs its golden files do not cover.  Each call sets up a fresh
!> run_params, V_ex and config the way the reference's initialisation does
!> (initialise.F90:153-298) and calls the reference's routine through its
!> procedure pointers.
module ref_shim

  use iso_c_binding
  use kinds, only : real64, int64
  use constants, only : array_int
  use shared_data, only : config, V_ex
  use derived_types, only : run_params
  use c_functions, only : f90_init_genrand, genrand
  use initialise, only : initialise_function_pointers, initial_setup
  use analytics, only : lattice_shells, radial_densities
  use bw_hamiltonian, only : pair_energy
  use random_site, only : pair_swap
  implicit none

contains

  subroutine make_setup(setup, lattice, n1, n2, n3, ns, nr)
    type(run_params), intent(out) :: setup
    integer(c_int32_t), intent(in) :: lattice, n1, n2, n3, ns, nr
    if (lattice == 1) then
      setup%lattice = 'fcc'
    else
      setup%lattice = 'bcc'
    end if
    setup%n_1 = n1
    setup%n_2 = n2
    setup%n_3 = n3
    setup%n_species = ns
    setup%interaction_range = nr
    setup%wc_range = 1
    setup%static_seed = .true.
    ! initialise_function_pointers checks that the concentrations sum to 1
    allocate(setup%species_numbers(ns), setup%species_concentrations(ns))
    setup%species_numbers = 0_int64
    setup%species_concentrations = 1.0_real64 / real(ns, real64)
    call initialise_function_pointers(setup)
    if (allocated(config)) deallocate(config)
    allocate(config(setup%n_basis, 2*n1, 2*n2, 2*n3))
    if (allocated(V_ex)) deallocate(V_ex)
    allocate(V_ex(ns, ns, nr))
  end subroutine make_setup

  subroutine load(setup, vex, cfg)
    type(run_params), intent(in) :: setup
    real(c_double), intent(in) :: vex(*)
    integer(c_int8_t), intent(in) :: cfg(*)
    integer :: nv, nc
    nv = size(V_ex)
    nc = size(config)
    V_ex = reshape(vex(1:nv), shape(V_ex))
    config = reshape(int(cfg(1:nc), array_int), shape(config))
  end subroutine load

  !> setup%full_energy(config) = total_energy (bw_hamiltonian.f90:58-81)
  function ref_total_energy(lattice, n1, n2, n3, ns, nr, vex, cfg) result(e) &
      bind(C, name="ref_total_energy")
    integer(c_int32_t), value :: lattice, n1, n2, n3, ns, nr
    real(c_double), intent(in) :: vex(*)
    integer(c_int8_t), intent(in) :: cfg(*)
    real(c_double) :: e
    type(run_params) :: setup
    call make_setup(setup, lattice, n1, n2, n3, ns, nr)
    call load(setup, vex, cfg)
    e = setup%full_energy(config)
  end function ref_total_energy

  !> setup%nbr_energy(config, 1, i, j, k) = *_energy_Rshells for every
  !> occupied site, in the reference's loop order (z, y, x), into e(n_atoms).
  subroutine ref_site_energies(lattice, n1, n2, n3, ns, nr, vex, cfg, e) &
      bind(C, name="ref_site_energies")
    integer(c_int32_t), value :: lattice, n1, n2, n3, ns, nr
    real(c_double), intent(in) :: vex(*)
    integer(c_int8_t), intent(in) :: cfg(*)
    real(c_double), intent(out) :: e(*)
    type(run_params) :: setup
    integer :: i, j, k, q
    call make_setup(setup, lattice, n1, n2, n3, ns, nr)
    call load(setup, vex, cfg)
    q = 0
    do k = 1, 2*n3
      do j = 1, 2*n2
        do i = 1, 2*n1
          if (config(1, i, j, k) == 0_array_int) cycle
          q = q + 1
          e(q) = setup%nbr_energy(config, 1, i, j, k)
        end do
      end do
    end do
  end subroutine ref_site_energies

  !> initial_setup (initialise.F90:434-617) after f90_init_genrand with a
  !> static seed for `rank`; species_numbers as given (length ns).
  subroutine ref_initial_setup(lattice, n1, n2, n3, ns, rank, species_numbers, cfg) &
      bind(C, name="ref_initial_setup")
    integer(c_int32_t), value :: lattice, n1, n2, n3, ns, rank
    integer(c_int64_t), intent(in) :: species_numbers(*)
    integer(c_int8_t), intent(out) :: cfg(*)
    type(run_params) :: setup
    integer(c_int) :: seed
    integer :: nc
    call make_setup(setup, lattice, n1, n2, n3, ns, 1)
    setup%species_numbers = species_numbers(1:ns)
    setup%species_concentrations = 0.0_real64
    seed = f90_init_genrand(0_c_int, int(rank, c_int), 0_c_int)
    call initial_setup(setup, config)
    nc = size(config)
    cfg(1:nc) = int(reshape(config, [nc]), c_int8_t)
  end subroutine ref_initial_setup

  !> `count` draws of setup%rdm_site (random_site.f90:127-193) after
  !> f90_init_genrand for `rank`; sites(4, count).
  subroutine ref_random_sites(lattice, n1, n2, n3, rank, count, sites) &
      bind(C, name="ref_random_sites")
    integer(c_int32_t), value :: lattice, n1, n2, n3, rank, count
    integer(c_int32_t), intent(out) :: sites(4, *)
    type(run_params) :: setup
    integer(c_int) :: seed
    integer :: q
    call make_setup(setup, lattice, n1, n2, n3, 1, 1)
    seed = f90_init_genrand(0_c_int, int(rank, c_int), 0_c_int)
    do q = 1, count
      sites(:, q) = setup%rdm_site()
    end do
  end subroutine ref_random_sites

  !> lattice_shells + radial_densities (analytics.f90:205-404).
  subroutine ref_radial_densities(lattice, n1, n2, n3, ns, wc, cfg, shells_out, rho) &
      bind(C, name="ref_radial_densities")
    integer(c_int32_t), value :: lattice, n1, n2, n3, ns, wc
    integer(c_int8_t), intent(in) :: cfg(*)
    real(c_double), intent(out) :: shells_out(*), rho(*)
    type(run_params) :: setup
    real(real64), allocatable :: shells(:)
    real(real64), allocatable :: r(:,:,:)
    real(c_double) :: dummy(1)
    integer :: nc
    call make_setup(setup, lattice, n1, n2, n3, ns, 1)
    setup%wc_range = wc
    dummy = 0.0_c_double
    nc = size(config)
    config = reshape(int(cfg(1:nc), array_int), shape(config))
    allocate(shells(wc))
    shells = 0.0_real64
    call lattice_shells(setup, shells, config)
    allocate(r(ns, ns, wc))
    r = radial_densities(setup, config, wc, shells)
    shells_out(1:wc) = shells
    rho(1:ns*ns*wc) = reshape(r, [ns*ns*wc])
  end subroutine ref_radial_densities

  !> The reference's serial Metropolis path for the CPU baseline: after
  !> f90_init_genrand for `rank`, n_sweeps sweeps of n_atoms Kawasaki trials
  !> at inverse temperature beta, each trial made of the reference's own
  !> routines in the order of monte_carlo_step_lattice
  !> (metropolis.F90:751-813: setup%rdm_site twice, equal species accepted,
  !> pair_energy, pair_swap, pair_energy, genrand against exp(-beta dE),
  !> pair_swap back on rejection), and setup%full_energy after every sweep
  !> as the reference's sampling does (metropolis.F90:361).  cfg in/out;
  !> returns the accepted count, the last energy in e_last and the seconds
  !> of the trial loops in seconds.
  function ref_metropolis_sweeps(lattice, n1, n2, n3, ns, nr, vex, cfg, beta, n_sweeps, rank, &
                                 e_last, seconds) result(acc) bind(C, name="ref_metropolis_sweeps")
    integer(c_int32_t), value :: lattice, n1, n2, n3, ns, nr, n_sweeps, rank
    real(c_double), value :: beta
    real(c_double), intent(in) :: vex(*)
    integer(c_int8_t), intent(inout) :: cfg(*)
    real(c_double), intent(out) :: e_last, seconds
    integer(c_int64_t) :: acc
    type(run_params) :: setup
    integer(c_int) :: seed
    integer :: sweep, t, nc, n_atoms
    integer, dimension(4) :: r1, r2
    real(real64) :: e_u, e_s, de
    integer(int64) :: c0, c1, rate
    call make_setup(setup, lattice, n1, n2, n3, ns, nr)
    call load(setup, vex, cfg)
    seed = f90_init_genrand(0_c_int, int(rank, c_int), 0_c_int)
    n_atoms = count(config /= 0_array_int)
    acc = 0_c_int64_t
    e_last = 0.0_c_double
    call system_clock(c0, rate)
    do sweep = 1, n_sweeps
      do t = 1, n_atoms
        r1 = setup%rdm_site()
        r2 = setup%rdm_site()
        if (config(r1(1), r1(2), r1(3), r1(4)) == config(r2(1), r2(2), r2(3), r2(4))) then
          acc = acc + 1
          cycle
        end if
        e_u = pair_energy(setup, config, r1, r2)
        call pair_swap(config, r1, r2)
        e_s = pair_energy(setup, config, r1, r2)
        de = e_s - e_u
        if (de < 0.0_real64) then
          acc = acc + 1
        else if (genrand() < exp(-beta * de)) then
          acc = acc + 1
        else
          call pair_swap(config, r1, r2)
        end if
      end do
      e_last = setup%full_energy(config)
    end do
    call system_clock(c1)
    seconds = real(c1 - c0, c_double) / real(rate, c_double)
    nc = size(config)
    cfg(1:nc) = int(reshape(config, [nc]), c_int8_t)
  end function ref_metropolis_sweeps

end module ref_shim
