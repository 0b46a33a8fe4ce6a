!> brawl_hip_adaptors -- procedures with the reference's own interfaces
!> (hamiltonian, neighbour, monte_carlo: derived_types.f90:373-486) that run
!> on the GPU through brawl_hip.  A BraWl host rebinds its procedure
!> pointers with one line each (initialise.F90:157-241, metropolis.F90:
!> 171-175):
!>
!>   call bwl_attach(setup)                  ! after V_ex is read
!>   setup%full_energy => bwl_full_energy
!>   setup%nbr_energy  => bwl_nbr_energy
!>   setup%mc_step     => bwl_mc_step_lattice ! or bwl_mc_step_nbr
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
            bwl_initial_setup_gpu

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

  !> initial_setup (initialise.F90:434-617) from the context's stream; the
  !> result is written to `config`.
  subroutine bwl_initial_setup_gpu(setup, config)
    class(run_params), intent(in) :: setup
    integer(array_int), dimension(:,:,:,:), intent(out) :: config
    integer(c_int64_t), allocatable, target :: sn(:)
    real(c_double), allocatable, target :: cc(:)
    type(c_ptr) :: psn, pcc
    psn = c_null_ptr
    pcc = c_null_ptr
    if (allocated(setup%species_numbers)) then
      allocate(sn(size(setup%species_numbers)))
      sn = setup%species_numbers
      psn = c_loc(sn)
    end if
    if (allocated(setup%species_concentrations)) then
      allocate(cc(size(setup%species_concentrations)))
      cc = setup%species_concentrations
      pcc = c_loc(cc)
    end if
    call bwl_check(bwl_initial_setup(ctx, psn, pcc), 'bwl_initial_setup')
    call bwl_check(bwl_get_config(ctx, config), 'bwl_get_config')
  end subroutine bwl_initial_setup_gpu

end module brawl_hip_adaptors
