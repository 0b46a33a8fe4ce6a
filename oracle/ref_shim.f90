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
  use constants, only : array_int, Ry_to_eV
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
    ! initialise_function_pointers checks that the concentrations sum to 1;
    ! the concentrations are allocated 0:ns with element 0 = 0, as
    ! read_inputs does (io.f90:242-246): initial_setup sums from element 0
    allocate(setup%species_numbers(ns), setup%species_concentrations(0:ns))
    setup%species_numbers = 0_int64
    setup%species_concentrations = 1.0_real64 / real(ns, real64)
    setup%species_concentrations(0) = 0.0_real64
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

  !> initial_setup on its concentrations path (initialise.F90:474-505):
  !> species_numbers left at 0, concentrations(1:ns) as given, element 0 = 0.
  subroutine ref_initial_setup_conc(lattice, n1, n2, n3, ns, rank, conc, cfg) &
      bind(C, name="ref_initial_setup_conc")
    integer(c_int32_t), value :: lattice, n1, n2, n3, ns, rank
    real(c_double), intent(in) :: conc(*)
    integer(c_int8_t), intent(out) :: cfg(*)
    type(run_params) :: setup
    integer(c_int) :: seed
    integer :: nc
    call make_setup(setup, lattice, n1, n2, n3, ns, 1)
    setup%species_numbers = 0_int64
    setup%species_concentrations(0) = 0.0_real64
    setup%species_concentrations(1:ns) = conc(1:ns)
    seed = f90_init_genrand(0_c_int, int(rank, c_int), 0_c_int)
    call initial_setup(setup, config)
    nc = size(config)
    cfg(1:nc) = int(reshape(config, [nc]), c_int8_t)
  end subroutine ref_initial_setup_conc

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

  !> `count` pairs (setup%rdm_site, setup%rdm_nbr of it) after
  !> f90_init_genrand for `rank`, as select_sites_nbr draws them
  !> (wang-landau.F90:1527-1532); sites(4, 2, count).
  subroutine ref_random_nbrs(lattice, n1, n2, n3, rank, count, sites) &
      bind(C, name="ref_random_nbrs")
    integer(c_int32_t), value :: lattice, n1, n2, n3, rank, count
    integer(c_int32_t), intent(out) :: sites(4, 2, *)
    type(run_params) :: setup
    integer(c_int) :: seed
    integer :: q
    call make_setup(setup, lattice, n1, n2, n3, 1, 1)
    seed = f90_init_genrand(0_c_int, int(rank, c_int), 0_c_int)
    do q = 1, count
      sites(:, 1, q) = setup%rdm_site()
      sites(:, 2, q) = setup%rdm_nbr(sites(:, 1, q))
    end do
  end subroutine ref_random_nbrs

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

  !> wang-landau.F90 bin_index (:515-523), restated here because the module
  !> needs an MPI module this image does not have.
  integer function wl_bin(energy, edges, bins)
    real(real64), intent(in) :: energy, edges(*)
    integer, intent(in) :: bins
    real(real64) :: bin_range
    bin_range = edges(bins + 1) - edges(1)
    wl_bin = int(((energy - edges(1))/(bin_range))*real(bins, real64)) + 1
  end function wl_bin

  !> n_calls `sweeps` calls (wang-landau.F90:539-632) of one walker after
  !> f90_init_genrand for `rank`: the loop of the reference with its own
  !> routines (setup%full_energy, setup%rdm_site / rdm_nbr as mc_select,
  !> pair_energy, pair_swap, genrand), without the window average over
  !> ranks and the radial-density and configuration stores (they draw no
  !> numbers).  The walker must start inside its window [start_idx,
  !> end_idx] (it then never leaves it).  cfg, lng (bins) and hist
  !> (end_idx-start_idx+1) in/out; returns the accepted count.
  function ref_wl_sweeps(lattice, n1, n2, n3, ns, nr, vex, cfg, rank, nbr_swap, mc_sweeps, &
                         n_calls, bins, edges, start_idx, end_idx, wl_f, lng, hist) result(acc) &
      bind(C, name="ref_wl_sweeps")
    integer(c_int32_t), value :: lattice, n1, n2, n3, ns, nr, rank, nbr_swap, mc_sweeps, n_calls, &
                                 bins, start_idx, end_idx
    real(c_double), value :: wl_f
    real(c_double), intent(in) :: vex(*), edges(*)
    integer(c_int8_t), intent(inout) :: cfg(*)
    real(c_double), intent(inout) :: lng(*), hist(*)
    integer(c_int64_t) :: acc
    type(run_params) :: setup
    integer(c_int) :: seed
    integer :: call_no, i, ibin, jbin, nc
    integer, dimension(4) :: rdm1, rdm2
    integer(array_int) :: site1, site2
    real(real64) :: e_swapped, e_unswapped, pair_unswapped, pair_swapped
    call make_setup(setup, lattice, n1, n2, n3, ns, nr)
    call load(setup, vex, cfg)
    seed = f90_init_genrand(0_c_int, int(rank, c_int), 0_c_int)
    acc = 0_c_int64_t
    do call_no = 1, n_calls
      e_unswapped = setup%full_energy(config)
      e_swapped = e_unswapped
      do i = 1, mc_sweeps*setup%n_atoms
        rdm1 = setup%rdm_site()
        if (nbr_swap /= 0) then
          rdm2 = setup%rdm_nbr(rdm1)
        else
          rdm2 = setup%rdm_site()
        end if
        site1 = config(rdm1(1), rdm1(2), rdm1(3), rdm1(4))
        site2 = config(rdm2(1), rdm2(2), rdm2(3), rdm2(4))
        e_swapped = e_unswapped
        pair_unswapped = pair_energy(setup, config, rdm1, rdm2)
        pair_swapped = pair_unswapped
        call pair_swap(config, rdm1, rdm2)
        if (site1 /= site2) then
          pair_swapped = pair_energy(setup, config, rdm1, rdm2)
          e_swapped = e_unswapped - pair_unswapped + pair_swapped
        end if
        ibin = wl_bin(e_unswapped, edges, bins)
        jbin = wl_bin(e_swapped, edges, bins)
        if (jbin > start_idx - 1 .and. jbin < end_idx + 1) then
          if (log(genrand()) .lt. (lng(ibin) - lng(jbin))) then
            acc = acc + 1
            e_unswapped = e_swapped
          else
            call pair_swap(config, rdm1, rdm2)
            jbin = ibin
          end if
        else
          jbin = ibin
          call pair_swap(config, rdm1, rdm2)
        end if
        if (MOD(i, INT(0.02_real64*REAL(setup%n_atoms))) == 0) then
          hist(jbin - start_idx + 1) = hist(jbin - start_idx + 1) + 1.0_real64
        end if
        lng(jbin) = lng(jbin) + wl_f
      end do
    end do
    nc = size(config)
    cfg(1:nc) = int(reshape(config, [nc]), c_int8_t)
  end function ref_wl_sweeps

  !> enter_energy_window (wang-landau.F90:643-741) of one walker after
  !> f90_init_genrand for `rank`, without the hand-off between the ranks of
  !> a window: the reference's loop with initial_setup every n_atoms*250
  !> iterations (species concentrations 1/ns), its own routines for the
  !> trials, stopping after max_trials trials.  min_e, max_e are the
  !> window's outer bin edges; energy_min, energy_max the REAL(4) bounds in
  !> meV/atom.  cfg in/out; returns the trials made.
  function ref_wl_enter(lattice, n1, n2, n3, ns, nr, vex, cfg, rank, min_e, max_e, energy_min, &
                        energy_max, max_trials) result(trials) bind(C, name="ref_wl_enter")
    integer(c_int32_t), value :: lattice, n1, n2, n3, ns, nr, rank
    real(c_double), value :: min_e, max_e
    real(c_float), value :: energy_min, energy_max
    integer(c_int64_t), value :: max_trials
    real(c_double), intent(in) :: vex(*)
    integer(c_int8_t), intent(inout) :: cfg(*)
    integer(c_int64_t) :: trials
    type(run_params) :: setup
    integer(c_int) :: seed
    integer :: i_steps, sweeps, nc
    integer, dimension(4) :: rdm1, rdm2
    integer(array_int) :: site1, site2
    real :: e_min4, e_max4
    real(real64) :: e_swapped, e_unswapped, pair_swapped, pair_unswapped, delta_e, target_energy, &
                    condition
    call make_setup(setup, lattice, n1, n2, n3, ns, nr)
    call load(setup, vex, cfg)
    seed = f90_init_genrand(0_c_int, int(rank, c_int), 0_c_int)
    e_min4 = energy_min
    e_max4 = energy_max
    target_energy = (min_e + max_e)/2.0_real64
    condition = ABS(max_e - min_e)*0.1_real64
    e_unswapped = setup%full_energy(config)
    trials = 0_c_int64_t
    i_steps = 0
    sweeps = 50
    do while (.true.)
      i_steps = i_steps + 1
      if (MOD(i_steps, setup%n_atoms*sweeps*5) == 0) then
        i_steps = 0
        call initial_setup(setup, config)
      end if
      if (e_unswapped < max_e - condition .and. e_unswapped > min_e + condition) then
        e_unswapped = setup%full_energy(config)
        if (e_unswapped < max_e - condition .and. e_unswapped > min_e + condition) then
        else
          cycle
        end if
        exit
      end if
      if (trials >= max_trials) exit
      trials = trials + 1
      rdm1 = setup%rdm_site()
      rdm2 = setup%rdm_site()
      site1 = config(rdm1(1), rdm1(2), rdm1(3), rdm1(4))
      site2 = config(rdm2(1), rdm2(2), rdm2(3), rdm2(4))
      if (site1 /= site2) then
        pair_unswapped = pair_energy(setup, config, rdm1, rdm2)
        call pair_swap(config, rdm1, rdm2)
        pair_swapped = pair_energy(setup, config, rdm1, rdm2)
        e_swapped = e_unswapped - pair_unswapped + pair_swapped
        delta_e = ((e_swapped-target_energy)**2-(e_unswapped-target_energy)**2)/ &
                  (2.0_real64*(0.0025*ABS(e_max4 - e_min4)*setup%n_atoms/(Ry_to_eV*1000))**2)
        if (log(genrand()) .lt. -delta_e) then
          e_unswapped = e_swapped
        else
          call pair_swap(config, rdm1, rdm2)
        end if
      end if
    end do
    nc = size(config)
    cfg(1:nc) = int(reshape(config, [nc]), c_int8_t)
  end function ref_wl_enter

  !> The nested-sampling loop (nested_sampling.f90:78-200) after
  !> f90_init_genrand for `rank`, with the reference's own routines
  !> (initial_setup with species_numbers, genrand, setup%full_energy,
  !> setup%rdm_site, pair_energy, pair_swap); the energies file, trajectory
  !> and screen output are left out.  n_walkers >= 4, so the first
  !> walk-length check follows a walk.  e_limits(n_iter) out, with the walk
  !> trials made and the seconds of the iteration loop; returns 0.
  function ref_nested_sampling(lattice, n1, n2, n3, ns, nr, vex, rank, species_numbers, n_walkers, &
                               n_steps, n_iter, e_limits, trials, seconds) result(status) &
      bind(C, name="ref_nested_sampling")
    integer(c_int32_t), value :: lattice, n1, n2, n3, ns, nr, rank, n_walkers, n_steps, n_iter
    real(c_double), intent(in) :: vex(*)
    integer(c_int64_t), intent(in) :: species_numbers(*)
    real(c_double), intent(out) :: e_limits(*)
    integer(c_int64_t), intent(out) :: trials
    real(c_double), intent(out) :: seconds
    integer(c_int32_t) :: status
    integer(int64) :: c0, c1, rate
    type(run_params) :: setup
    integer(c_int) :: seed
    integer(array_int), allocatable :: walkers(:,:,:,:,:)
    real(real64), allocatable :: walker_energies(:)
    integer :: i_walker, i_iter, i_step, extra_steps, n_acc, n_at, irnd
    integer, dimension(1) :: i_max
    integer, dimension(4) :: rdm1, rdm2
    integer(array_int) :: site1, site2
    real(real64) :: rnde, rnd, ener_limit, e_unswapped, e_swapped, delta_e
    call make_setup(setup, lattice, n1, n2, n3, ns, nr)
    call load_vex(vex)
    setup%species_numbers = species_numbers(1:ns)
    setup%species_concentrations = 0.0_real64
    seed = f90_init_genrand(0_c_int, int(rank, c_int), 0_c_int)
    allocate(walkers(setup%n_basis, 2*n1, 2*n2, 2*n3, n_walkers), walker_energies(n_walkers))
    n_at = setup%n_1*setup%n_2*setup%n_3*setup%n_basis*setup%n_species
    do i_walker = 1, n_walkers
      call initial_setup(setup, config)
      walkers(:,:,:,:,i_walker) = config
      rnde = genrand()
      walker_energies(i_walker) = setup%full_energy(walkers(:,:,:,:,i_walker)) + rnde*1e-8
    end do
    extra_steps = 0
    n_acc = 0
    trials = 0_c_int64_t
    call system_clock(c0, rate)
    do i_iter = 1, n_iter
      i_max = maxloc(walker_energies)
      ener_limit = walker_energies(i_max(1))
      e_limits(i_iter) = ener_limit
      if (mod(i_iter, int(n_walkers/2.0)) == 0) then
        if ((n_acc < n_at*0.05) .and. (extra_steps < n_steps*100)) then
          extra_steps = extra_steps + int(n_steps)
        end if
      end if
      rnd = genrand()
      irnd = ceiling(rnd*n_walkers)
      walkers(:,:,:,:,i_max(1)) = walkers(:,:,:,:,irnd)
      walker_energies(i_max(1)) = walker_energies(irnd)
      n_acc = 0
      trials = trials + int(n_steps + extra_steps, c_int64_t)
      do i_step = 1, n_steps + extra_steps
        rdm1 = setup%rdm_site()
        site1 = walkers(rdm1(1), rdm1(2), rdm1(3), rdm1(4), i_max(1))
        do
          rdm2 = setup%rdm_site()
          site2 = walkers(rdm2(1), rdm2(2), rdm2(3), rdm2(4), i_max(1))
          if (site1 /= site2) exit
        end do
        e_unswapped = pair_energy(setup, walkers(:,:,:,:,i_max(1)), rdm1, rdm2)
        call pair_swap(walkers(:,:,:,:,i_max(1)), rdm1, rdm2)
        e_swapped = pair_energy(setup, walkers(:,:,:,:,i_max(1)), rdm1, rdm2)
        delta_e = e_swapped - e_unswapped
        if (walker_energies(i_max(1)) + delta_e < ener_limit) then
          walker_energies(i_max(1)) = walker_energies(i_max(1)) + delta_e
          n_acc = n_acc + 1
        else
          call pair_swap(walkers(:,:,:,:,i_max(1)), rdm1, rdm2)
        end if
      end do
    end do
    call system_clock(c1)
    seconds = real(c1 - c0, c_double) / real(rate, c_double)
    status = 0
  end function ref_nested_sampling

  subroutine load_vex(vex)
    real(c_double), intent(in) :: vex(*)
    integer :: nv
    nv = size(V_ex)
    V_ex = reshape(vex(1:nv), shape(V_ex))
  end subroutine load_vex

end module ref_shim
