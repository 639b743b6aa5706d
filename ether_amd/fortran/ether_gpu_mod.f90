! ether_gpu_mod.f90 - iso_c_binding interface to libether_gpu.so
! (include/ether_gpu.h).  This is the thin layer between the Fortran driver
! and the gfx950 kernels: every routine below is a bind(C) declaration of one
! C-ABI entry point; arrays are passed by address, sizes by value.
module ether_gpu
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: ETHER_OK = 0
   integer(c_int), parameter :: ETHER_ERR_NODEV = 5
   integer(c_int), parameter :: ETHER_MODEL_ISING = 0, ETHER_MODEL_XYZ = 1
   integer(c_int), parameter :: ETHER_INIT_FRESH = 1, ETHER_INIT_RANDOM = 2
   integer(c_int), parameter :: ETHER_COMPAT_PT_ETOT = 1
   integer(c_int), parameter :: ETHER_COMPAT_SPLIT_OVRR = 2
   integer(c_int), parameter :: ETHER_COMPAT_MEASURE_PASS = 4
   integer(c_int), parameter :: ETHER_COMPAT_CHAIN_REPEATS = 8
   integer(c_int), parameter :: ETHER_MOM_COUNT = 8

   interface
      integer(c_int) function ether_build_neighbours(abc, n_basis, basis_xyz, sc, open_bc, &
            n_shells, shell_d, tol, capacity, nn_count, nn_idx) bind(C, name="ether_build_neighbours")
         import :: c_int, c_double
         real(c_double), intent(in) :: abc(*), basis_xyz(*), shell_d(*)
         integer(c_int), value :: n_basis, n_shells, capacity
         integer(c_int), intent(in) :: sc(*), open_bc(*)
         real(c_double), value :: tol
         integer(c_int), intent(out) :: nn_count(*), nn_idx(*)
      end function

      integer(c_int) function ether_build_neighbours_device(abc, n_basis, basis_xyz, sc, open_bc, &
            n_shells, shell_d, tol, capacity, nn_count, nn_idx, colour, n_colours) &
            bind(C, name="ether_build_neighbours_device")
         import :: c_int, c_double
         real(c_double), intent(in) :: abc(*), basis_xyz(*), shell_d(*)
         integer(c_int), value :: n_basis, n_shells, capacity
         integer(c_int), intent(in) :: sc(*), open_bc(*)
         real(c_double), value :: tol
         integer(c_int), intent(out) :: nn_count(*), nn_idx(*), colour(*)
         integer(c_int), intent(out) :: n_colours
      end function

      integer(c_int) function ether_colour_lattice(abc, n_basis, basis_xyz, sc, open_bc, n_shells, &
            shell_d, tol, colour, n_colours) bind(C, name="ether_colour_lattice")
         import :: c_int, c_double
         real(c_double), intent(in) :: abc(*), basis_xyz(*), shell_d(*)
         integer(c_int), value :: n_basis, n_shells
         integer(c_int), intent(in) :: sc(*), open_bc(*)
         real(c_double), value :: tol
         integer(c_int), intent(out) :: colour(*)
         integer(c_int), intent(out) :: n_colours
      end function

      integer(c_int) function ether_start_state(seed, model, angle, n_basis, file_index, S, phi) &
            bind(C, name="ether_start_state")
         import :: c_int, c_int64_t, c_double
         integer(c_int64_t), value :: seed
         integer(c_int), value :: model, angle, n_basis
         integer(c_int), intent(in) :: file_index(*)
         real(c_double), intent(out) :: S(*), phi(*)
      end function

      integer(c_int) function ether_ladder_train(ctx, n_temps, temps, kb, first_step, train_mcs, &
            windows, validate_mcs, coa, ovrr_interval, ovrr_passes, pt_interval, learn, placement, &
            mlp_seed, swap_count, T_out, report) bind(C, name="ether_ladder_train")
         import :: c_int, c_int64_t, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: n_temps, first_step, train_mcs, windows, validate_mcs, coa
         integer(c_int), value :: ovrr_interval, pt_interval, learn, placement
         real(c_double), intent(in) :: temps(*)
         real(c_double), value :: kb, ovrr_passes
         integer(c_int64_t), value :: mlp_seed
         integer(c_int), intent(inout) :: swap_count
         real(c_double), intent(out) :: T_out(*), report(*)
      end function

      integer(c_int) function ether_ctx_create(model, n_sites, n_replicas, seed, ctx) &
            bind(C, name="ether_ctx_create")
         import :: c_int, c_int64_t, c_ptr
         integer(c_int), value :: model, n_sites, n_replicas
         integer(c_int64_t), value :: seed
         type(c_ptr), intent(out) :: ctx
      end function

      subroutine ether_ctx_destroy(ctx) bind(C, name="ether_ctx_destroy")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine

      integer(c_int) function ether_last_error(ctx, buf, len) bind(C, name="ether_last_error")
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_int), value :: len
      end function

      integer(c_int) function ether_set_lattice(ctx, n_species, species, stagger, n_shells, &
            capacity, nn_count, nn_idx, colour) bind(C, name="ether_set_lattice")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: n_species, n_shells, capacity
         integer(c_int), intent(in) :: species(*), nn_count(*), nn_idx(*)
         real(c_double), intent(in) :: stagger(*)
         type(c_ptr), value :: colour
      end function

      integer(c_int) function ether_set_exchange(ctx, j_exc, spin) bind(C, name="ether_set_exchange")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: j_exc(*), spin(*)
      end function

      integer(c_int) function ether_set_zeeman(ctx, on, h3, g_factor, mu_b) bind(C, name="ether_set_zeeman")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: on
         real(c_double), intent(in) :: h3(3)
         real(c_double), value :: g_factor, mu_b
      end function

      integer(c_int) function ether_set_sia(ctx, on, sia) bind(C, name="ether_set_sia")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: on
         real(c_double), intent(in) :: sia(*)
      end function

      integer(c_int) function ether_set_proposal(ctx, angle, dphi) bind(C, name="ether_set_proposal")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: angle
         real(c_double), value :: dphi
      end function

      integer(c_int) function ether_set_compat(ctx, flags) bind(C, name="ether_set_compat")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: flags
      end function

      integer(c_int) function ether_commit(ctx) bind(C, name="ether_commit")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function ether_set_betas(ctx, beta) bind(C, name="ether_set_betas")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: beta(*)
      end function

      integer(c_int) function ether_set_ladders(ctx, n_groups, group_start) bind(C, name="ether_set_ladders")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: n_groups
         integer(c_int), intent(in) :: group_start(*)
      end function

      integer(c_int) function ether_init_spins(ctx, mode) bind(C, name="ether_init_spins")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: mode
      end function

      integer(c_int) function ether_set_spins(ctx, replica, S, phi) bind(C, name="ether_set_spins")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: replica
         real(c_double), intent(in) :: S(*), phi(*)
      end function

      integer(c_int) function ether_get_spins(ctx, replica, S, phi) bind(C, name="ether_get_spins")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: replica
         real(c_double), intent(out) :: S(*), phi(*)
      end function

      integer(c_int) function ether_reset_moments(ctx) bind(C, name="ether_reset_moments")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function ether_set_graphs(ctx, on) bind(C, name="ether_set_graphs")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: on
      end function

      integer(c_int) function ether_set_launch(ctx, sites_per_wave, xcd_map) bind(C, name="ether_set_launch")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: sites_per_wave, xcd_map
      end function

      integer(c_int) function ether_set_lanes(ctx, lanes_per_site) bind(C, name="ether_set_lanes")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: lanes_per_site
      end function

      integer(c_int) function ether_build_spiral_partners(n_sites, pos, dir, dis, latency_deg, &
            capacity, partners, count) bind(C, name="ether_build_spiral_partners")
         import :: c_int, c_int64_t, c_double
         integer(c_int), value :: n_sites, capacity
         real(c_double), intent(in) :: pos(*), dir(*)
         real(c_double), value :: dis, latency_deg
         integer(c_int), intent(out) :: partners(*)
         integer(c_int64_t), intent(out) :: count
      end function

      integer(c_int) function ether_set_spiral(ctx, capacity, partners) bind(C, name="ether_set_spiral")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: capacity
         integer(c_int), intent(in) :: partners(*)
      end function

      integer(c_int) function ether_read_spiral(ctx, moments, psi) bind(C, name="ether_read_spiral")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: moments(*)
         type(c_ptr), value :: psi
      end function

      integer(c_int) function ether_read_moments(ctx, moments) bind(C, name="ether_read_moments")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: moments(*)
      end function

      integer(c_int) function ether_run(ctx, first_step, n_steps, mcs_eq, coa, ovrr_interval, &
            ovrr_passes, pt_interval, swap_count) bind(C, name="ether_run")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: first_step, n_steps, mcs_eq, coa, ovrr_interval, pt_interval
         real(c_double), value :: ovrr_passes
         integer(c_int), intent(inout) :: swap_count
      end function

      integer(c_int) function ether_autotune(ctx, reps, sites_per_wave, xcd_map, table_ms) &
            bind(C, name="ether_autotune")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: reps
         integer(c_int), intent(out) :: sites_per_wave, xcd_map
         type(c_ptr), value :: table_ms
      end function

      integer(c_int) function ether_batch_create(ctxs, n, batch) bind(C, name="ether_batch_create")
         import :: c_int, c_ptr
         type(c_ptr), intent(in) :: ctxs(*)
         integer(c_int), value :: n
         type(c_ptr), intent(out) :: batch
      end function

      integer(c_int) function ether_batch_run(batch, first_step, n_steps, mcs_eq, coa, ovrr_interval, &
            ovrr_passes, pt_interval, swap_count) bind(C, name="ether_batch_run")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: batch
         integer(c_int), value :: first_step, n_steps, mcs_eq, coa, ovrr_interval, pt_interval
         real(c_double), value :: ovrr_passes
         integer(c_int), intent(inout) :: swap_count
      end function

      subroutine ether_batch_destroy(batch) bind(C, name="ether_batch_destroy")
         import :: c_ptr
         type(c_ptr), value :: batch
      end subroutine

      integer(c_int) function ether_batch_set_graphs(batch, on) bind(C, name="ether_batch_set_graphs")
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
         integer(c_int), value :: on
      end function

      integer(c_int) function ether_batch_get_resident(batch, on) bind(C, name="ether_batch_get_resident")
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
         integer(c_int), intent(out) :: on
      end function

      integer(c_int) function ether_set_resident(ctx, mode) bind(C, name="ether_set_resident")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: mode
      end function

      integer(c_int) function ether_get_resident(ctx, kind) bind(C, name="ether_get_resident")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: kind
      end function

      integer(c_int) function ether_pt_state(ctx, slot_of, acc, att, trips) bind(C, name="ether_pt_state")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: slot_of(*)
         integer(c_int64_t), intent(out) :: acc(*), att(*), trips(*)
      end function

      integer(c_int) function ether_synchronize(ctx) bind(C, name="ether_synchronize")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
   end interface

contains

   ! Message of the last failure (ctx may be c_null_ptr for context-free calls)
   function ether_message(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=512) :: msg
      character(kind=c_char) :: buf(512)
      integer :: i, rc
      rc = ether_last_error(ctx, buf, 512_c_int)
      msg = ' '
      do i = 1, 512
         if (buf(i) == c_null_char) exit
         msg(i:i) = buf(i)
      end do
   end function

end module ether_gpu
