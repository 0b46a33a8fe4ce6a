!> brawl_hip_adaptors -- procedures with the reference's own interfaces
!> (hamiltonian, neighbour, rand_site, rand_neighbour, monte_carlo:
!> derived_types.f90:373-486; wl_params%mc_select: :349) that run on the
!> GPU through brawl_hip.  A BraWl host rebinds its procedure pointers with
!> one line each (initialise.F90:157-241, metropolis.F90:171-175,
!> wang-landau.F90:110-114):
!>
!>   call bwl_attach(setup)                  ! after V_ex is read
!>   setup%full_energy => bwl_full_energy
!>   setup%nbr_energy  => bwl_nbr_energy
!>   setup%rdm_site    => bwl_rdm_site
!>   setup%rdm_nbr     => bwl_rdm_nbr
!>   setup%mc_step     => bwl_mc_step_lattice ! or bwl_mc_step_nbr
!>   wl_setup%mc_select => bwl_mc_select      ! or bwl_mc_select_nbr
!>
!> This module uses the reference's modules (kinds, constants,
!> derived_types, shared_data), so it is compiled inside a BraWl build.
!>
!> Random numbers.  The reference's MT19937 state is private to
!> mt19937ar.c; the MC-step adaptors draw from the device context's own
!> copy of the same generator.  bwl_seed_stream seeds it exactly as
!> f90_init_genrand does for a static seed, so a host whose draws all go
!> through the adaptors (bwl_initial_setup_gpu, the MC steps) reproduces the
!> reference's stream; bwl_genrand_get_state / bwl_genrand_set_state move
!> the state when the host also draws.
module brawl_hip_adaptors

  use iso_c_binding, only : c_ptr, c_null_ptr, c_associated, c_int, c_int8_t, c_int32_t, &
                            c_int64_t, c_double, c_loc
  use kinds, only : real64, int64
  use constants, only : array_int
  use derived_types, only : run_params
  use shared_data, only : V_ex
  use brawl_hip
  implicit none

  private
  public :: bwl_attach, bwl_detach, bwl_context, bwl_seed_stream, bwl_full_energy, &
            bwl_nbr_energy, bwl_mc_step_lattice, bwl_mc_step_nbr, bwl_metropolis_sweeps, &
            bwl_initial_setup_gpu, bwl_rdm_site, bwl_rdm_nbr, bwl_mc_select, &
            bwl_mc_select_nbr

  type(c_ptr), save :: ctx = c_null_ptr

contains

  !> Create the device context for `setup` (lattice, cells, species,
  !> interaction range) and upload V_ex (shared_data.f90:37).
  subroutine bwl_attach(setup)
    class(run_params), intent(in) :: setup
    type(bwl_lattice_desc) :: d
    select case (trim(setup%lattice))
    case ('fcc')
      d%lattice = BWL_FCC
    case ('bcc')
      d%lattice = BWL_BCC
    case default
      stop 'brawl_hip: only fcc and bcc lattices run on the GPU'
    end select
    d%n1 = setup%n_1
    d%n2 = setup%n_2
    d%n3 = setup%n_3
    d%n_species = setup%n_species
    d%n_shells = setup%interaction_range
    call bwl_detach()
    call bwl_check(bwl_create(d, V_ex, ctx), 'bwl_create')
  end subroutine bwl_attach

  subroutine bwl_detach()
    if (c_associated(ctx)) call bwl_check(bwl_destroy(ctx), 'bwl_destroy')
    ctx = c_null_ptr
  end subroutine bwl_detach

  !> The device context (for the sweep-level and batched C-ABI calls).
  function bwl_context() result(c)
    type(c_ptr) :: c
    c = ctx
  end function bwl_context

  !> Seed the context's stream as f90_init_genrand does for a static seed
  !> (110179 + 11*rank, mt19937ar.c:136-139).
  subroutine bwl_seed_stream(rank)
    integer, intent(in) :: rank
    call bwl_check(bwl_genrand_seed(ctx, int(110179 + 11*rank, c_int32_t)), 'bwl_genrand_seed')
  end subroutine bwl_seed_stream

  subroutine upload(config)
    integer(array_int), dimension(:,:,:,:), intent(in) :: config
    call bwl_check(bwl_set_config(ctx, config), 'bwl_set_config')
  end subroutine upload

  !> `hamiltonian` (bound as setup%full_energy): total energy from the
  !> device's integer pair counts, contracted in FP64.
  function bwl_full_energy(setup, config) result(energy)
    class(run_params), intent(in) :: setup
    integer(array_int), dimension(:,:,:,:), intent(in) :: config
    real(real64) :: energy
    call upload(config)
    call bwl_check(bwl_total_energy(ctx, energy), 'bwl_total_energy')
  end function bwl_full_energy

  !> `neighbour` (bound as setup%nbr_energy): one site's energy in the
  !> reference's FP64 summation order (bit-identical).
  function bwl_nbr_energy(setup, config, site_b, site_i, site_j, site_k) result(energy)
    class(run_params), intent(in) :: setup
    integer(array_int), dimension(:,:,:,:), intent(in) :: config
    integer, intent(in) :: site_b, site_i, site_j, site_k
    real(real64) :: energy
    call upload(config)
    call bwl_check(bwl_site_energy(ctx, int(site_b, c_int32_t), int(site_i, c_int32_t), &
                                   int(site_j, c_int32_t), int(site_k, c_int32_t), energy), &
                   'bwl_site_energy')
  end function bwl_nbr_energy

  integer function mc_steps(config, beta, nbr, n) result(accept)
    integer(array_int), dimension(:,:,:,:) :: config
    real(real64), intent(in) :: beta
    integer(c_int32_t), intent(in) :: nbr
    integer(c_int64_t), intent(in) :: n
    integer(c_int64_t) :: acc
    call upload(config)
    call bwl_check(bwl_mc_steps(ctx, beta, nbr, n, 0_c_int64_t, c_null_ptr, acc), 'bwl_mc_steps')
    call bwl_check(bwl_get_config(ctx, config), 'bwl_get_config')
    accept = int(acc)
  end function mc_steps

  !> `monte_carlo` (bound as setup%mc_step): one lattice-wide swap trial
  !> (monte_carlo_step_lattice, metropolis.F90:751-813) on the device.
  function bwl_mc_step_lattice(setup, config, beta) result(accept)
    class(run_params), intent(in) :: setup
    integer(array_int), dimension(:,:,:,:) :: config
    real(real64), intent(in) :: beta
    integer :: accept
    accept = mc_steps(config, beta, 0_c_int32_t, 1_c_int64_t)
  end function bwl_mc_step_lattice

  !> `monte_carlo`: one nearest-neighbour swap trial
  !> (monte_carlo_step_nbr, metropolis.F90:829-891).
  function bwl_mc_step_nbr(setup, config, beta) result(accept)
    class(run_params), intent(in) :: setup
    integer(array_int), dimension(:,:,:,:) :: config
    real(real64), intent(in) :: beta
    integer :: accept
    accept = mc_steps(config, beta, 1_c_int32_t, 1_c_int64_t)
  end function bwl_mc_step_nbr

  !> The sweep-level drop-in for the trial loop of metropolis.F90:350-354:
  !> n_trials sequential trials in one launch, accepted count returned.
  function bwl_metropolis_sweeps(config, beta, n_trials, nbr_swap) result(accepted)
    integer(array_int), dimension(:,:,:,:) :: config
    real(real64), intent(in) :: beta
    integer(int64), intent(in) :: n_trials
    logical, intent(in) :: nbr_swap
    integer(int64) :: accepted
    integer(c_int32_t) :: nbr
    nbr = 0
    if (nbr_swap) nbr = 1
    accepted = int(mc_steps(config, beta, nbr, int(n_trials, c_int64_t)), int64)
  end function bwl_metropolis_sweeps

  !> `rand_site` (bound as setup%rdm_site; bcc_random_site /
  !> fcc_random_site, random_site.f90:127-193): one site drawn from the
  !> context's stream as the reference draws it.
  function bwl_rdm_site(setup) result(site)
    class(run_params), intent(in) :: setup
    integer, dimension(4) :: site
    integer(c_int32_t) :: s(4, 1)
    call bwl_check(bwl_random_site(ctx, 1_c_int32_t, s), 'bwl_random_site')
    site = int(s(:, 1))
  end function bwl_rdm_site

  !> `rand_neighbour` (bound as setup%rdm_nbr; *_random_nbr,
  !> random_site.f90:152-225): a shell-1 neighbour of `site`, one draw.
  function bwl_rdm_nbr(setup, site) result(nbr)
    class(run_params), intent(in) :: setup
    integer, dimension(4), intent(in) :: site
    integer, dimension(4) :: nbr
    integer(c_int32_t) :: s(4), t(4)
    s = int(site, c_int32_t)
    call bwl_check(bwl_random_nbr(ctx, s, t), 'bwl_random_nbr')
    nbr = int(t)
  end function bwl_rdm_nbr

  !> wl_params%mc_select for lattice-wide swaps (select_sites,
  !> wang-landau.F90:1534-1539): two random sites.
  subroutine bwl_mc_select(rdm1, rdm2)
    integer, dimension(4) :: rdm1, rdm2
    integer(c_int32_t) :: s(4, 2, 1)
    call bwl_check(bwl_select_sites(ctx, 0_c_int32_t, 1_c_int32_t, s), 'bwl_select_sites')
    rdm1 = int(s(:, 1, 1))
    rdm2 = int(s(:, 2, 1))
  end subroutine bwl_mc_select

  !> wl_params%mc_select for nearest-neighbour swaps (select_sites_nbr,
  !> wang-landau.F90:1527-1532): a random site and a shell-1 neighbour.
  subroutine bwl_mc_select_nbr(rdm1, rdm2)
    integer, dimension(4) :: rdm1, rdm2
    integer(c_int32_t) :: s(4, 2, 1)
    call bwl_check(bwl_select_sites(ctx, 1_c_int32_t, 1_c_int32_t, s), 'bwl_select_sites')
    rdm1 = int(s(:, 1, 1))
    rdm2 = int(s(:, 2, 1))
  end subroutine bwl_mc_select_nbr
  !> initial_setup (initialise.F90:434-617) from the context's stream; the
  !> result is written to `config`.  The concentrations are taken from
  !> elements 1..n_species of setup%species_concentrations, which the
  !> reference allocates 0:n_species with element 0 = 0 (io.f90:242).  On
  !> the species-numbers path the reference also stores the REAL(4)
  !> concentrations count/n_atoms in setup (initialise.F90:468-472); so
  !> does this adaptor.
  subroutine bwl_initial_setup_gpu(setup, config)
    class(run_params), intent(inout) :: setup
    integer(array_int), dimension(:,:,:,:), intent(out) :: config
    integer(c_int64_t), allocatable, target :: sn(:)
    real(c_double), allocatable, target :: cc(:)
    type(c_ptr) :: psn, pcc
    integer :: ns, ub, i
    ns = setup%n_species
    psn = c_null_ptr
    pcc = c_null_ptr
    if (allocated(setup%species_numbers)) then
      allocate(sn(ns))
      sn = setup%species_numbers(lbound(setup%species_numbers, 1):lbound(setup%species_numbers, 1)+ns-1)
      psn = c_loc(sn)
    end if
    if (allocated(setup%species_concentrations)) then
      ub = ubound(setup%species_concentrations, 1)
      allocate(cc(ns))
      cc = setup%species_concentrations(ub-ns+1:ub)
      pcc = c_loc(cc)
    end if
    call bwl_check(bwl_initial_setup(ctx, psn, pcc), 'bwl_initial_setup')
    call bwl_check(bwl_get_config(ctx, config), 'bwl_get_config')
    if (allocated(setup%species_numbers) .and. allocated(setup%species_concentrations)) then
      if (sum(setup%species_numbers) == setup%n_atoms) then
        ub = ubound(setup%species_concentrations, 1)
        do i = 1, ns
          setup%species_concentrations(ub-ns+i) = real(real(setup%species_numbers(i)) / real(setup%n_atoms), c_double)
        end do
      end if
    end if
  end subroutine bwl_initial_setup_gpu

end module brawl_hip_adaptors
