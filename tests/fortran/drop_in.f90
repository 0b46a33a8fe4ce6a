!> drop_in -- BraWl host code that rebinds its energy and MC-step
!> procedure pointers to the GPU adaptors (brawl_hip_adaptors) and checks
!> them against the reference's own routines on the same inputs.  Built as
!> a shared library (oracle/_ref/libdrop_in.so) whose C entry point
!> tests/test_gpu_fortran.py calls in-process:
!>   int drop_in_run(int lattice, int n, int S, int R, int n_trials,
!>                   const char *out_file)
!> It writes `key value` lines to out_file and returns the first error
!> status of the GPU library (0 if none).  Two more entry points check the
!> set-up and proposal adaptors:
!>   int drop_in_setup(int lattice, int n, int S, int rank, int use_numbers,
!>                     const int64_t *numbers, const double *conc,
!>                     int8_t *cfg_gpu, int8_t *cfg_ref,
!>                     double *conc_gpu, double *conc_ref)
!>   int drop_in_draws(int lattice, int n, int rank, int count,
!>                     int *mismatches)   /* 4 counts */
module drop_in

  use iso_c_binding
  use kinds, only : real64, int64
  use constants, only : array_int
  use shared_data, only : config, V_ex
  use derived_types, only : run_params, wl_params
  use c_functions, only : f90_init_genrand
  use initialise, only : initialise_function_pointers, initial_setup
  use brawl_hip
  use brawl_hip_adaptors
  implicit none

contains

  function drop_in_run(lattice, n, ns, nr, n_trials, path) result(status) bind(C, name="drop_in_run")
  integer(c_int), value :: lattice, n, ns, nr, n_trials
  character(kind=c_char), intent(in) :: path(*)
  integer(c_int) :: status
  type(run_params) :: setup
  character(len=512) :: out_file
  integer :: i, j, k, a, b, s, u, q, mism, seed
  integer(array_int), allocatable :: cfg_ref(:,:,:,:), cfg_gpu(:,:,:,:)
  real(real64) :: e_ref, e_gpu, es_ref, es_gpu
  integer :: acc
  integer(int64) :: acc_total
  type(bwl_anneal_params) :: ap
  type(bwl_run_stats) :: st
  real(c_double) :: e_T(1), c_T(1), a_T(1)
  integer(c_int) :: arc

  if (lattice == 1) then
    setup%lattice = 'fcc'
  else
    setup%lattice = 'bcc'
  end if
  out_file = ''
  do i = 1, len(out_file)
    if (path(i) == c_null_char) exit
    out_file(i:i) = path(i)
  end do
  bwl_stop_on_error = .false.
  bwl_first_error = BWL_OK
  if (allocated(config)) deallocate(config)
  if (allocated(V_ex)) deallocate(V_ex)

  ! the reference's set-up of a static-seed run on rank 0
  setup%n_1 = n
  setup%n_2 = n
  setup%n_3 = n
  setup%n_species = ns
  setup%interaction_range = nr
  setup%wc_range = 2
  setup%static_seed = .true.
  ! species numbers first: initialise_function_pointers checks them
  if (trim(setup%lattice) == 'fcc') then
    setup%n_atoms = 4*n*n*n
  else
    setup%n_atoms = 2*n*n*n
  end if
  ! concentrations allocated 0:ns with element 0 = 0, as read_inputs does
  ! (io.f90:242-246); initial_setup sums them from element 0
  allocate(setup%species_numbers(ns), setup%species_concentrations(0:ns))
  setup%species_numbers = int(setup%n_atoms / ns, int64)
  setup%species_numbers(1) = setup%species_numbers(1) + int(setup%n_atoms - ns*(setup%n_atoms/ns), int64)
  setup%species_concentrations = 0.0_real64
  call initialise_function_pointers(setup)
  allocate(config(setup%n_basis, 2*n, 2*n, 2*n))
  config = 0_array_int
  allocate(V_ex(ns, ns, nr))
  ! a symmetric, zero-mean-free deterministic interaction table (Ry)
  do s = 1, nr
    do b = 1, ns
      do a = 1, ns
        V_ex(a, b, s) = 1.0e-3_real64 * sin(real(a*b + 3*s, real64)) / real(s, real64)
      end do
    end do
  end do
  seed = f90_init_genrand(0_c_int, 0_c_int, 0_c_int)
  call initial_setup(setup, config)

  open(newunit=u, file=trim(out_file), status='replace', action='write')

  ! 1. the reference's energies
  e_ref = setup%full_energy(config)
  ! 2. rebind to the GPU and evaluate the same configuration
  call bwl_attach(setup)
  setup%full_energy => bwl_full_energy
  e_gpu = setup%full_energy(config)
  write(u, '(a,1x,es26.17e3)') 'e_ref', e_ref
  write(u, '(a,1x,es26.17e3)') 'e_gpu', e_gpu

  ! site energies through the `neighbour` pointer, both bindings
  mism = 0
  q = 0
  do k = 1, 2*n
    do j = 1, 2*n
      do i = 1, 2*n
        if (config(1, i, j, k) == 0_array_int) cycle
        q = q + 1
        if (mod(q, 7) /= 1) cycle
        call initialise_function_pointers(setup)
        es_ref = setup%nbr_energy(config, 1, i, j, k)
        setup%nbr_energy => bwl_nbr_energy
        es_gpu = setup%nbr_energy(config, 1, i, j, k)
        if (transfer(es_ref, 0_int64) /= transfer(es_gpu, 0_int64)) mism = mism + 1
      end do
    end do
  end do
  write(u, '(a,1x,i0)') 'site_energy_checked', (q + 6) / 7
  write(u, '(a,1x,i0)') 'site_energy_mismatches', mism

  ! 3. initial_setup from the device stream seeded as the reference's
  allocate(cfg_ref(size(config,1), size(config,2), size(config,3), size(config,4)))
  allocate(cfg_gpu(size(config,1), size(config,2), size(config,3), size(config,4)))
  cfg_ref = config
  call bwl_seed_stream(0)
  call bwl_initial_setup_gpu(setup, cfg_gpu)
  write(u, '(a,1x,i0)') 'initial_setup_mismatches', count(cfg_ref /= cfg_gpu)

  ! 4. MC steps through the rebound `monte_carlo` pointer, continuing the
  ! stream; the test replays the same trials on the oracle
  setup%mc_step => bwl_mc_step_lattice
  acc_total = 0
  do q = 1, n_trials
    acc = setup%mc_step(cfg_gpu, 1.0_real64 / (600.0_real64 * 8.167333262e-5_real64 / 13.605693122990_real64))
    acc_total = acc_total + acc
  end do
  write(u, '(a,1x,i0)') 'mc_accepted', acc_total
  write(u, '(a)', advance='no') 'mc_config'
  do k = 1, 2*n
    do j = 1, 2*n
      do i = 1, 2*n
        write(u, '(1x,i0)', advance='no') cfg_gpu(1, i, j, k)
      end do
    end do
  end do
  write(u, *)

  ! 5. the whole temperature step through the parallel sampler, called as a
  ! host calls it directly through the C-ABI (one sampled sweep: <E> is the
  ! energy of the configuration it leaves)
  ap = bwl_anneal_params(T=600.0_c_double, delta_T=0.0_c_double, T_steps=1, n_sweeps=1, &
                         burn_in_sweeps=3, asro_every=0, seed=7_c_int64_t)
  ! (a box too small for the sampler's rounds is refused with a status)
  call bwl_check(bwl_set_config(bwl_context(), cfg_gpu), 'bwl_set_config')
  arc = bwl_metropolis_anneal(bwl_context(), ap, e_T, c_T, a_T, st)
  write(u, '(a,1x,i0)') 'anneal_status', arc
  call bwl_check(bwl_get_config(bwl_context(), cfg_gpu), 'bwl_get_config')
  write(u, '(a,1x,es26.17e3)') 'anneal_E', e_T(1)
  write(u, '(a,1x,i0)') 'anneal_trials', st%trials
  write(u, '(a)', advance='no') 'anneal_config'
  do k = 1, 2*n
    do j = 1, 2*n
      do i = 1, 2*n
        write(u, '(1x,i0)', advance='no') cfg_gpu(1, i, j, k)
      end do
    end do
  end do
  write(u, *)
  close(u)
  call bwl_detach()
  status = bwl_first_error
  end function drop_in_run

  !> A setup as read_inputs leaves it (io.f90:242-246: concentrations
  !> 0:ns, element 0 = 0); the caller fills in numbers or concentrations.
  subroutine make_setup(setup, lattice, n, ns)
    type(run_params), intent(out) :: setup
    integer(c_int), intent(in) :: lattice, n, ns
    if (lattice == 1) then
      setup%lattice = 'fcc'
      setup%n_atoms = 4*n*n*n
    else
      setup%lattice = 'bcc'
      setup%n_atoms = 2*n*n*n
    end if
    setup%n_1 = n
    setup%n_2 = n
    setup%n_3 = n
    setup%n_species = ns
    setup%interaction_range = 1
    setup%wc_range = 2
    setup%static_seed = .true.
    allocate(setup%species_numbers(ns), setup%species_concentrations(0:ns))
    setup%species_numbers = 0_int64
    setup%species_concentrations = 0.0_real64
  end subroutine make_setup

  !> initial_setup of `rank`'s static-seed stream both ways: the
  !> reference's routine and the GPU adaptor, each on its own setup, on the
  !> species-numbers path (use_numbers /= 0) or the concentrations path.
  !> Writes both configurations and both setups' concentrations(1:ns)
  !> afterwards (the reference stores count/n_atoms there on the numbers
  !> path).
  function drop_in_setup(lattice, n, ns, rank, use_numbers, numbers, conc, cfg_gpu, cfg_ref, &
                         conc_gpu, conc_ref) result(status) bind(C, name="drop_in_setup")
    integer(c_int), value :: lattice, n, ns, rank, use_numbers
    integer(c_int64_t), intent(in) :: numbers(*)
    real(c_double), intent(in) :: conc(*)
    integer(c_int8_t), intent(out) :: cfg_gpu(*), cfg_ref(*)
    real(c_double), intent(out) :: conc_gpu(*), conc_ref(*)
    integer(c_int) :: status
    type(run_params) :: s_ref, s_gpu
    integer(array_int), allocatable :: c_gpu(:,:,:,:)
    integer :: seed, nc
    bwl_stop_on_error = .false.
    bwl_first_error = BWL_OK
    call make_setup(s_ref, lattice, n, ns)
    call make_setup(s_gpu, lattice, n, ns)
    if (use_numbers /= 0) then
      s_ref%species_numbers = numbers(1:ns)
      s_gpu%species_numbers = numbers(1:ns)
    else
      s_ref%species_concentrations(1:ns) = conc(1:ns)
      s_gpu%species_concentrations(1:ns) = conc(1:ns)
    end if
    call initialise_function_pointers(s_ref)
    call initialise_function_pointers(s_gpu)
    if (allocated(config)) deallocate(config)
    if (allocated(V_ex)) deallocate(V_ex)
    allocate(config(s_ref%n_basis, 2*n, 2*n, 2*n))
    allocate(V_ex(ns, ns, 1))
    V_ex = 0.0_real64
    config = 0_array_int
    ! the reference
    seed = f90_init_genrand(0_c_int, int(rank, c_int), 0_c_int)
    call initial_setup(s_ref, config)
    nc = size(config)
    cfg_ref(1:nc) = int(reshape(config, [nc]), c_int8_t)
    conc_ref(1:ns) = s_ref%species_concentrations(1:ns)
    ! the adaptor, from the same stream on the device
    allocate(c_gpu(s_gpu%n_basis, 2*n, 2*n, 2*n))
    call bwl_attach(s_gpu)
    call bwl_seed_stream(rank)
    call bwl_initial_setup_gpu(s_gpu, c_gpu)
    cfg_gpu(1:nc) = int(reshape(c_gpu, [nc]), c_int8_t)
    conc_gpu(1:ns) = s_gpu%species_concentrations(1:ns)
    call bwl_detach()
    status = bwl_first_error
  end function drop_in_setup

  !> The proposal adaptors against the reference's own draws from the same
  !> stream: setup%rdm_site, setup%rdm_nbr (of the site just drawn) and
  !> wl_params%mc_select both ways (select_sites: two rdm_site; and
  !> select_sites_nbr: rdm_site then rdm_nbr, wang-landau.F90:1527-1539),
  !> `count` draws each.  mism(1:4) = mismatching draws of rdm_site,
  !> rdm_nbr, mc_select, mc_select_nbr.
  function drop_in_draws(lattice, n, rank, count, mism) result(status) bind(C, name="drop_in_draws")
    integer(c_int), value :: lattice, n, rank, count
    integer(c_int), intent(out) :: mism(4)
    integer(c_int) :: status
    type(run_params) :: s_ref, s_gpu
    type(wl_params) :: wl
    integer, allocatable :: ref(:,:,:)
    integer, dimension(4) :: a, b
    integer :: seed, q, mode
    bwl_stop_on_error = .false.
    bwl_first_error = BWL_OK
    mism = 0
    call make_setup(s_ref, lattice, n, 2)
    s_ref%species_concentrations(1:2) = 0.5_real64
    call make_setup(s_gpu, lattice, n, 2)
    s_gpu%species_concentrations(1:2) = 0.5_real64
    call initialise_function_pointers(s_ref)
    call initialise_function_pointers(s_gpu)
    if (allocated(V_ex)) deallocate(V_ex)
    allocate(V_ex(2, 2, 1))
    V_ex = 0.0_real64
    call bwl_attach(s_gpu)
    s_gpu%rdm_site => bwl_rdm_site
    s_gpu%rdm_nbr => bwl_rdm_nbr
    allocate(ref(4, 2, count))
    do mode = 1, 4
      ! the reference's draws
      seed = f90_init_genrand(0_c_int, int(rank, c_int), 0_c_int)
      do q = 1, count
        ref(:, 1, q) = s_ref%rdm_site()
        if (mode == 2 .or. mode == 4) then
          ref(:, 2, q) = s_ref%rdm_nbr(ref(:, 1, q))
        else if (mode == 3) then
          ref(:, 2, q) = s_ref%rdm_site()
        end if
      end do
      ! the rebound pointers, same stream on the device
      call bwl_seed_stream(rank)
      if (mode == 3) wl%mc_select => bwl_mc_select
      if (mode == 4) wl%mc_select => bwl_mc_select_nbr
      do q = 1, count
        if (mode <= 2) then
          a = s_gpu%rdm_site()
          b = ref(:, 2, q)
          if (mode == 2) b = s_gpu%rdm_nbr(a)
        else
          call wl%mc_select(a, b)
        end if
        if (any(a /= ref(:, 1, q))) then
          mism(mode) = mism(mode) + 1
        else if (mode >= 2 .and. any(b /= ref(:, 2, q))) then
          mism(mode) = mism(mode) + 1
        end if
      end do
    end do
    call bwl_detach()
    status = bwl_first_error
  end function drop_in_draws

end module drop_in
