!> brawl_hip -- ISO_C_BINDING interface to libbrawl_hip.so, the gfx950
!> lattice-sampling library (include/brawl_hip.h), for a Fortran host.
!>
!> Every C-ABI entry point has an explicit bind(C) interface here with the
!> C types of the header; derived types mirror the header's structs.  Arrays
!> are passed by address (assumed-size dummies), so a Fortran caller passes
!> its own config(1,2n1,2n2,2n3) int8 array and V_ex(S,S,R) real64 array
!> as they are (contiguous, column-major): the layouts at the boundary are
!> the reference's.  bwl_check turns a nonzero status into the reference's
!> error convention (message, then stop).
module brawl_hip

  use iso_c_binding, only : c_int, c_int8_t, c_int32_t, c_int64_t, c_double, c_float, &
                            c_ptr, c_char, c_null_char, c_associated, c_f_pointer, c_null_ptr
  implicit none

  integer(c_int), parameter :: BWL_OK = 0, BWL_ERR_ARG = -1, BWL_ERR_UNSUPPORTED = -2, &
                               BWL_ERR_DEVICE = -3, BWL_ERR_NOMEM = -4, BWL_ERR_STATE = -5
  integer(c_int32_t), parameter :: BWL_FCC = 1, BWL_BCC = 2

  !> bwl_check follows the reference's convention (message, then stop)
  !> unless a host clears this; it then records the first failing status in
  !> bwl_first_error and returns.
  logical, save :: bwl_stop_on_error = .true.
  integer(c_int), save :: bwl_first_error = BWL_OK

  type, bind(C) :: bwl_lattice_desc
    integer(c_int32_t) :: lattice, n1, n2, n3, n_species, n_shells
  end type bwl_lattice_desc

  type, bind(C) :: bwl_anneal_params
    real(c_double) :: T, delta_T
    integer(c_int32_t) :: T_steps, n_sweeps, burn_in_sweeps, asro_every
    integer(c_int64_t) :: seed
  end type bwl_anneal_params

  type, bind(C) :: bwl_run_stats
    integer(c_int64_t) :: trials, accepted
    real(c_double) :: kernel_ms
    real(c_double) :: e_drift_abs, e_drift_rel
  end type bwl_run_stats

  type, bind(C) :: bwl_ns_params
    integer(c_int32_t) :: n_walkers, n_steps, n_iter, reserved
  end type bwl_ns_params

  type, bind(C) :: bwl_wl_params
    integer(c_int32_t) :: bins, num_windows, walkers_per_window
    real(c_float) :: bin_overlap, energy_min, energy_max
    integer(c_int32_t) :: nbr_swap, mc_sweeps
    integer(c_int32_t) :: rank_offset   ! uint32_t in C
  end type bwl_wl_params

  interface

    type(c_ptr) function bwl_last_error() bind(C, name="bwl_last_error")
      import :: c_ptr
    end function

    integer(c_int) function bwl_device_count(count) bind(C, name="bwl_device_count")
      import :: c_int, c_int32_t
      integer(c_int32_t), intent(out) :: count
    end function

    integer(c_int) function bwl_create(desc, v_ex, ctx) bind(C, name="bwl_create")
      import :: c_int, c_double, c_ptr, bwl_lattice_desc
      type(bwl_lattice_desc), intent(in) :: desc
      real(c_double), intent(in) :: v_ex(*)
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function bwl_destroy(ctx) bind(C, name="bwl_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function bwl_n_atoms(ctx, n_atoms) bind(C, name="bwl_n_atoms")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: n_atoms
    end function

    integer(c_int) function bwl_set_config(ctx, config) bind(C, name="bwl_set_config")
      import :: c_int, c_ptr, c_int8_t
      type(c_ptr), value :: ctx
      integer(c_int8_t), intent(in) :: config(*)
    end function

    integer(c_int) function bwl_get_config(ctx, config) bind(C, name="bwl_get_config")
      import :: c_int, c_ptr, c_int8_t
      type(c_ptr), value :: ctx
      integer(c_int8_t), intent(out) :: config(*)
    end function

    integer(c_int) function bwl_pair_histogram(ctx, hist) bind(C, name="bwl_pair_histogram")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: hist(*)
    end function

    integer(c_int) function bwl_total_energy(ctx, energy) bind(C, name="bwl_total_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: energy
    end function

    integer(c_int) function bwl_site_energy(ctx, site_b, site_i, site_j, site_k, energy) &
        bind(C, name="bwl_site_energy")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: site_b, site_i, site_j, site_k
      real(c_double), intent(out) :: energy
    end function

    integer(c_int) function bwl_pair_energy(ctx, site1, site2, energy) bind(C, name="bwl_pair_energy")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: site1(4), site2(4)
      real(c_double), intent(out) :: energy
    end function

    integer(c_int) function bwl_genrand_seed(ctx, seed) bind(C, name="bwl_genrand_seed")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: seed   ! uint32_t in C
    end function

    integer(c_int) function bwl_genrand_get_state(ctx, mt, mti) bind(C, name="bwl_genrand_get_state")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: mt(624), mti
    end function

    integer(c_int) function bwl_genrand_set_state(ctx, mt, mti) bind(C, name="bwl_genrand_set_state")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: mt(624)
      integer(c_int32_t), value :: mti
    end function

    integer(c_int) function bwl_random_site(ctx, n, sites) bind(C, name="bwl_random_site")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(out) :: sites(4, *)
    end function

    integer(c_int) function bwl_random_nbr(ctx, site, nbr) bind(C, name="bwl_random_nbr")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: site(4)
      integer(c_int32_t), intent(out) :: nbr(4)
    end function

    integer(c_int) function bwl_select_sites(ctx, nbr_swap, n, sites) bind(C, name="bwl_select_sites")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nbr_swap, n
      integer(c_int32_t), intent(out) :: sites(4, 2, *)
    end function

    integer(c_int) function bwl_initial_setup(ctx, species_numbers, concentrations) &
        bind(C, name="bwl_initial_setup")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: species_numbers, concentrations   ! int64_t*, double* (may be null)
    end function

    integer(c_int) function bwl_mc_steps(ctx, beta, nbr_swap, n_trials, sample_every, energies, &
                                         accepted) bind(C, name="bwl_mc_steps")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), value :: beta
      integer(c_int32_t), value :: nbr_swap
      integer(c_int64_t), value :: n_trials, sample_every
      type(c_ptr), value :: energies                    ! double* (may be null)
      integer(c_int64_t), intent(out) :: accepted
    end function

    integer(c_int) function bwl_metropolis_anneal(ctx, p, energies_of_T, C_of_T, acceptance_of_T, &
                                                  stats) bind(C, name="bwl_metropolis_anneal")
      import :: c_int, c_ptr, c_double, bwl_anneal_params, bwl_run_stats
      type(c_ptr), value :: ctx
      type(bwl_anneal_params), intent(in) :: p
      real(c_double), intent(out) :: energies_of_T(*), C_of_T(*), acceptance_of_T(*)
      type(bwl_run_stats), intent(out) :: stats
    end function

    integer(c_int) function bwl_metropolis_anneal_rho(ctx, p, energies_of_T, C_of_T, acceptance_of_T, &
                                                      rho_of_T, stats) bind(C, name="bwl_metropolis_anneal_rho")
      import :: c_int, c_ptr, c_double, bwl_anneal_params, bwl_run_stats
      type(c_ptr), value :: ctx
      type(bwl_anneal_params), intent(in) :: p
      real(c_double), intent(out) :: energies_of_T(*), C_of_T(*), acceptance_of_T(*), rho_of_T(*)
      type(bwl_run_stats), intent(out) :: stats
    end function

    integer(c_int) function bwl_metropolis_anneal_batch(n, ctxs, params, energies_of_T, C_of_T, &
                                                        acceptance_of_T, stats, makespan_ms) &
        bind(C, name="bwl_metropolis_anneal_batch")
      import :: c_int, c_int32_t, c_ptr, c_double, bwl_anneal_params, bwl_run_stats
      integer(c_int32_t), value :: n
      type(c_ptr), intent(in) :: ctxs(*)
      type(bwl_anneal_params), intent(in) :: params(*)
      type(c_ptr), intent(in) :: energies_of_T(*), C_of_T(*), acceptance_of_T(*)
      type(bwl_run_stats), intent(out) :: stats(*)
      real(c_double), intent(out) :: makespan_ms
    end function

    integer(c_int) function bwl_lattice_shells(ctx, wc_range, shells) bind(C, name="bwl_lattice_shells")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: wc_range
      real(c_double), intent(out) :: shells(*)
    end function

    integer(c_int) function bwl_radial_densities(ctx, wc_range, rho) bind(C, name="bwl_radial_densities")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: wc_range
      real(c_double), intent(out) :: rho(*)
    end function

    integer(c_int) function bwl_nested_sampling(ctx, p, species_numbers, concentrations, e_limits, &
                                                stats) bind(C, name="bwl_nested_sampling")
      import :: c_int, c_ptr, c_double, bwl_ns_params, bwl_run_stats
      type(c_ptr), value :: ctx
      type(bwl_ns_params), intent(in) :: p
      type(c_ptr), value :: species_numbers, concentrations
      real(c_double), intent(out) :: e_limits(*)
      type(bwl_run_stats), intent(out) :: stats
    end function

    integer(c_int) function bwl_nested_sampling_batch(n, ctxs, params, species_numbers, &
                                                      concentrations, e_limits, stats, makespan_ms) &
        bind(C, name="bwl_nested_sampling_batch")
      import :: c_int, c_int32_t, c_ptr, c_double, bwl_ns_params, bwl_run_stats
      integer(c_int32_t), value :: n
      type(c_ptr), intent(in) :: ctxs(*)
      type(bwl_ns_params), intent(in) :: params(*)
      type(c_ptr), intent(in) :: species_numbers(*), concentrations(*), e_limits(*)
      type(bwl_run_stats), intent(out) :: stats(*)
      real(c_double), intent(out) :: makespan_ms
    end function

    integer(c_int) function bwl_wl_create(desc, v_ex, p, wl) bind(C, name="bwl_wl_create")
      import :: c_int, c_double, c_ptr, bwl_lattice_desc, bwl_wl_params
      type(bwl_lattice_desc), intent(in) :: desc
      real(c_double), intent(in) :: v_ex(*)
      type(bwl_wl_params), intent(in) :: p
      type(c_ptr), intent(out) :: wl
    end function

    integer(c_int) function bwl_wl_destroy(wl) bind(C, name="bwl_wl_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: wl
    end function

    integer(c_int) function bwl_wl_n_walkers(wl, n_walkers) bind(C, name="bwl_wl_n_walkers")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: wl
      integer(c_int32_t), intent(out) :: n_walkers
    end function

    integer(c_int) function bwl_wl_windows(wl, windows, bin_edges) bind(C, name="bwl_wl_windows")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: wl
      integer(c_int32_t), intent(out) :: windows(*)
      real(c_double), intent(out) :: bin_edges(*)
    end function

    integer(c_int) function bwl_wl_hist_stride(wl, stride) bind(C, name="bwl_wl_hist_stride")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: wl
      integer(c_int32_t), intent(out) :: stride
    end function

    integer(c_int) function bwl_wl_initial_setup(wl, species_numbers, concentrations) &
        bind(C, name="bwl_wl_initial_setup")
      import :: c_int, c_ptr
      type(c_ptr), value :: wl, species_numbers, concentrations
    end function

    integer(c_int) function bwl_wl_set_walker(wl, walker, config) bind(C, name="bwl_wl_set_walker")
      import :: c_int, c_ptr, c_int32_t, c_int8_t
      type(c_ptr), value :: wl
      integer(c_int32_t), value :: walker
      integer(c_int8_t), intent(in) :: config(*)
    end function

    integer(c_int) function bwl_wl_get_walker(wl, walker, config) bind(C, name="bwl_wl_get_walker")
      import :: c_int, c_ptr, c_int32_t, c_int8_t
      type(c_ptr), value :: wl
      integer(c_int32_t), value :: walker
      integer(c_int8_t), intent(out) :: config(*)
    end function

    integer(c_int) function bwl_wl_genrand_get_state(wl, walker, mt, mti) &
        bind(C, name="bwl_wl_genrand_get_state")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: wl
      integer(c_int32_t), value :: walker
      integer(c_int32_t), intent(out) :: mt(624), mti
    end function

    integer(c_int) function bwl_wl_genrand_set_state(wl, walker, mt, mti) &
        bind(C, name="bwl_wl_genrand_set_state")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: wl
      integer(c_int32_t), value :: walker, mti
      integer(c_int32_t), intent(in) :: mt(624)
    end function

    integer(c_int) function bwl_wl_get_lng(wl, lng, hist) bind(C, name="bwl_wl_get_lng")
      import :: c_int, c_ptr
      type(c_ptr), value :: wl, lng, hist            ! double* (may be null)
    end function

    integer(c_int) function bwl_wl_set_lng(wl, lng, hist) bind(C, name="bwl_wl_set_lng")
      import :: c_int, c_ptr
      type(c_ptr), value :: wl, lng, hist
    end function

    integer(c_int) function bwl_wl_enter_windows(wl, max_trials, trials, entered) &
        bind(C, name="bwl_wl_enter_windows")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: wl
      integer(c_int64_t), value :: max_trials
      integer(c_int64_t), intent(out) :: trials(*)
      integer(c_int32_t), intent(out) :: entered(*)
    end function

    integer(c_int) function bwl_wl_sweeps(wl, wl_f, average, accepted, kernel_ms) &
        bind(C, name="bwl_wl_sweeps")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: wl
      real(c_double), value :: wl_f
      integer(c_int32_t), value :: average
      integer(c_int64_t), intent(out) :: accepted(*)
      real(c_double), intent(out) :: kernel_ms
    end function

    ! batched monte_carlo steps: one chain per context, one launch
    integer(c_int) function bwl_mc_steps_batch(n, ctxs, beta, nbr_swap, n_trials, sample_every, &
                                               energies, accepted, kernel_ms) &
        bind(C, name="bwl_mc_steps_batch")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      integer(c_int32_t), value :: n
      type(c_ptr), intent(in) :: ctxs(*)
      real(c_double), intent(in) :: beta(*)
      integer(c_int32_t), value :: nbr_swap
      integer(c_int64_t), value :: n_trials, sample_every
      type(c_ptr), intent(in) :: energies(*)      ! double* per chain (may be null)
      integer(c_int64_t), intent(out) :: accepted(*)
      real(c_double), intent(out) :: kernel_ms
    end function

    ! the parallel sampler's lattice layout in LDS (no reference counterpart)
    integer(c_int) function bwl_set_layout(ctx, layout) bind(C, name="bwl_set_layout")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: layout
    end function

    ! threads per chain of the parallel sampler (no reference counterpart)
    integer(c_int) function bwl_set_threads(ctx, threads) bind(C, name="bwl_set_threads")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: threads
    end function

    integer(c_int) function bwl_layouts(ctx, mask) bind(C, name="bwl_layouts")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: mask
    end function

    integer(c_int) function bwl_layout_info(ctx, layout, threads, lds_bytes) &
        bind(C, name="bwl_layout_info")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: layout
      integer(c_int32_t), intent(out) :: threads
      integer(c_int64_t), intent(out) :: lds_bytes
    end function

    ! replica_exchange (wang-landau.F90:1392-1501) over every walker
    integer(c_int) function bwl_wl_replica_exchange(wl, partner, exchanged, kernel_ms) &
        bind(C, name="bwl_wl_replica_exchange")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: wl
      integer(c_int32_t), intent(out) :: partner(*)
      integer(c_int32_t), intent(out) :: exchanged(*)
      real(c_double), intent(out) :: kernel_ms
    end function

    integer(c_int) function c_strlen(s) bind(C, name="strlen")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function

  end interface

contains

  !> The message of the last failure on this thread, as a Fortran string.
  function bwl_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = bwl_last_error()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = c_strlen(p)
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function bwl_error_message

  !> The reference's error convention: print the message and stop (or,
  !> with bwl_stop_on_error cleared, record the status and return).
  subroutine bwl_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= BWL_OK) then
      print '(a,a,a,i0,a,a)', 'brawl_hip: ', where, ' failed (', rc, '): ', bwl_error_message()
      if (bwl_stop_on_error) stop 'GPU lattice library error'
      if (bwl_first_error == BWL_OK) bwl_first_error = rc
    end if
  end subroutine bwl_check

end module brawl_hip
