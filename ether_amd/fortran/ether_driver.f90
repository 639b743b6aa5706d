! ether_driver.f90 - Fortran driver of the MI355X sweep engine.
!
! Reads the current directory exactly as the reference does: in.ether
! (tags and defaults of src/read_input.f90:61-148, src/defaults.f90:25-56),
! structure.vasp (src/read_structure.f90:51-106), j_exchange
! (src/j_values.f90:39-87), single_ion_anisotropy (src/get_sia_values.f90:38-62)
! and staggered (src/get_staggered_info.f90:42-76).  It builds the supercell
! in the reference's site order (src/generate_supercell.f90:104-132), then
! runs the temperature / repeat / MCS schedule of src/ether.f90:79-214 on the
! GPU through the C-ABI (module ether_gpu): every temperature x repeat chain
! is a device replica and all of them advance together.  Results go to
! energy.dat, magnetization.dat and spiral_state_parameters.dat in the
! reference's format (src/write_output_files.f90:27-98).  Any failure prints a remark and stops
! with a nonzero status, like the reference's terminate (src/terminate.f90).
module ether_driver_input
   use iso_c_binding
   implicit none
   integer, parameter :: dp = 8

   ! tags with the reference's defaults
   character(len=5) :: model = 'ising'
   integer :: tmcs = 5000, tmcs_eq = 3000, sc(3) = (/2, 2, 2/), nrepeat = 10, ovrr_mcs = 0
   integer :: seed_value = 1992, exchange_interval = 11, nbd_capacity = 20, spiral_capacity = 5
   real(dp) :: ht = 100.0_dp, lt = 10.0_dp, tint = 5.0_dp, dphi = 0.0_dp, coa = 10.0_dp
   real(dp) :: h(3) = 0.0_dp, g_factor = 2.0_dp, j_para = 1.0_dp, ovrr_para = 0.0_dp
   real(dp) :: nbdfc = 1.0d-5, ss_dis = 4.0_dp, ss_lat = 2.0_dp, ss_direc(3) = (/0.0_dp, 0.0_dp, 1.0_dp/)
   logical :: stg = .false., angle = .false., zeeman = .false., sia = .false., para = .false.
   logical :: ovrr = .false., pt = .false., cb = .false., ssp = .false.
   ! gpu.ether: engine settings the reference has no tags for
   integer :: gpu_autotune = 0, gpu_spw = 0, gpu_xcd = 1, gpu_lanes = 0, gpu_compat = 0, gpu_chunk = 0
   logical :: gpu_graphs = .true., gpu_learned = .false.
   integer :: gpu_train_mcs = 0, gpu_windows = 4
   character :: bc(3) = (/'c', 'c', 'c'/)
   real(dp), allocatable :: spin(:)
   character(len=2), allocatable :: species_incl(:)
   logical :: species_given = .false.

   ! structure.vasp
   integer :: nspecies = 0, natoms = 0
   character(len=2), allocatable :: species_name(:)
   integer, allocatable :: ions(:)
   real(dp) :: lat(3, 3)                 ! lat(:, a) = lattice vector a
   real(dp), allocatable :: xyz(:, :)    ! (3, natoms) Cartesian

   real(dp), parameter :: pi = 4.0_dp*atan(1.0_dp)

contains

   subroutine terminate(remark)
      character(len=*), intent(in) :: remark
      write (6, '(a)') ''
      write (6, '(a)') ' ==> ETHER (GPU) stopped: '//trim(remark)
      write (6, '(a)') ''
      error stop 1
   end subroutine

   function lower(s) result(o)
      character(len=*), intent(in) :: s
      character(len=len(s)) :: o
      integer :: i, c
      o = s
      do i = 1, len(s)
         c = iachar(s(i:i))
         if (c >= iachar('A') .and. c <= iachar('Z')) o(i:i) = achar(c + 32)
      end do
   end function

   ! first letter upper, second lower (species symbols)
   function sym(s) result(o)
      character(len=*), intent(in) :: s
      character(len=2) :: o
      character(len=len(s)) :: t
      integer :: c
      t = adjustl(s)
      o = t(1:min(2, len(t)))
      c = iachar(o(1:1))
      if (c >= iachar('a') .and. c <= iachar('z')) o(1:1) = achar(c - 32)
      c = iachar(o(2:2))
      if (c >= iachar('A') .and. c <= iachar('Z')) o(2:2) = achar(c + 32)
   end function

   integer function count_tokens(line) result(n)
      character(len=*), intent(in) :: line
      integer :: i
      logical :: inword
      n = 0
      inword = .false.
      do i = 1, len_trim(line)
         if (line(i:i) /= ' ' .and. line(i:i) /= ',' .and. line(i:i) /= achar(9)) then
            if (.not. inword) n = n + 1
            inword = .true.
         else
            inword = .false.
         end if
      end do
   end function

   ! get_tot_species (src/get_tot_species.f90:30-57)
   subroutine read_species_count()
      character(len=400) :: line
      integer :: i, u, ios
      logical :: ok
      inquire (file='structure.vasp', exist=ok)
      if (.not. ok) call terminate("Required file 'structure.vasp' is not present")
      open (newunit=u, file='structure.vasp', status='old', action='read')
      do i = 1, 6
         read (u, '(a)', iostat=ios) line
      end do
      if (ios /= 0) call terminate('Incorrect structure file')
      nspecies = count_tokens(line)
      close (u)
      allocate (spin(nspecies))
      spin = 1.0_dp
   end subroutine

   ! read_input with process_file semantics (src/read_input.f90:41-150,
   ! src/process_file.f90:45-69)
   subroutine read_input()
      character(len=400) :: line, key, value
      integer :: u, ios, i, j, k, n
      logical :: ok
      inquire (file='in.ether', exist=ok)
      if (.not. ok) call terminate("Required file 'in.ether' is not present")
      open (newunit=u, file='in.ether', status='old', action='read')
      do
         read (u, '(a)', iostat=ios) line
         if (ios /= 0) exit
         do i = 1, len(line)
            if (line(i:i) == achar(9)) line(i:i) = ' '
         end do
         i = scan(line, '!#')
         if (i > 0) line = line(:i - 1)
         if (len_trim(line) == 0) cycle
         line = lower(line)
         j = index(line, '=')
         if (j <= 0) call terminate("Missing '=' in the given line: "//trim(adjustl(line)))
         key = adjustl(line(:j - 1))
         value = line(j + 1:)
         select case (trim(key))
         case ('model')
            read (value, *, iostat=ios) model
            if (model /= 'ising' .and. model /= 'xyz') call terminate('unknown model '//trim(value))
         case ('mcs')
            read (value, *, iostat=ios) tmcs, tmcs_eq
            if (tmcs <= tmcs_eq) call terminate('TOTAL MCS  < EQUILIBRATIONS MCS steps')
         case ('temp')
            read (value, *, iostat=ios) ht, lt, tint
         case ('spin')
            read (value, *, iostat=ios) (spin(k), k=1, nspecies)
         case ('species')
            n = count_tokens(value)
            if (allocated(species_incl)) deallocate (species_incl)
            allocate (species_incl(n))
            read (value, *, iostat=ios) (species_incl(k), k=1, n)
            do k = 1, n
               species_incl(k) = sym(species_incl(k))
            end do
            species_given = .true.
         case ('sc')
            read (value, *, iostat=ios) sc(1:3)
         case ('stg')
            read (value, *, iostat=ios) stg
         case ('bc')
            read (value, *, iostat=ios) bc(1:3)
         case ('repeat')
            read (value, *, iostat=ios) nrepeat
         case ('angle')
            read (value, *, iostat=ios) angle, dphi
            dphi = pi/180.0_dp*dphi
         case ('coa')
            read (value, *, iostat=ios) coa
         case ('zeeman')
            read (value, *, iostat=ios) zeeman, h(1:3)
         case ('g_factor')
            read (value, *, iostat=ios) g_factor
         case ('sia')
            read (value, *, iostat=ios) sia
         case ('para')
            read (value, *, iostat=ios) para, j_para
            j_para = abs(j_para)
            if (j_para == 0.0_dp .and. para) &
               call terminate('parameter value is set ON and value cannot be zero')
         case ('ovrr')
            read (value, *, iostat=ios) ovrr, ovrr_para, ovrr_mcs
            ovrr_para = abs(ovrr_para)
         case ('seed')
            read (value, *, iostat=ios) seed_value
         case ('nbdfc')
            read (value, *, iostat=ios) nbdfc
         case ('pt')
            read (value, *, iostat=ios) pt, exchange_interval
         case ('cb')
            read (value, *, iostat=ios) cb
         case ('ss_dis')
            read (value, *, iostat=ios) ss_dis
         case ('ss_lat')
            read (value, *, iostat=ios) ss_lat
         case ('ss_direc')
            read (value, *, iostat=ios) ss_direc(1:3)
         case ('ssp')
            read (value, *, iostat=ios) ssp
         case ('spiral_capacity')
            read (value, *, iostat=ios) spiral_capacity
         case ('nbd_capacity')
            read (value, *, iostat=ios) nbd_capacity
         case default
            call terminate('Found unknown/missing information: '//trim(adjustl(line)))
         end select
         if (ios /= 0) call terminate('cannot read the value of tag '//trim(key))
      end do
      close (u)
   end subroutine

   ! gpu.ether (optional): same syntax as in.ether; kept in its own file because
   ! the reference's parser stops on unknown keys (src/read_input.f90:144-146)
   subroutine read_gpu_input()
      character(len=400) :: line, key, value
      character(len=16) :: word
      integer :: u, ios, i, j
      logical :: ok
      inquire (file='gpu.ether', exist=ok)
      if (.not. ok) return
      open (newunit=u, file='gpu.ether', status='old', action='read')
      do
         read (u, '(a)', iostat=ios) line
         if (ios /= 0) exit
         do i = 1, len(line)
            if (line(i:i) == achar(9)) line(i:i) = ' '
         end do
         i = scan(line, '!#')
         if (i > 0) line = line(:i - 1)
         if (len_trim(line) == 0) cycle
         line = lower(line)
         j = index(line, '=')
         if (j <= 0) call terminate("Missing '=' in the given line: "//trim(adjustl(line)))
         key = adjustl(line(:j - 1))
         value = line(j + 1:)
         ios = 0
         select case (trim(key))
         case ('ladder')
            read (value, *, iostat=ios) word
            if (trim(word) /= 'linear' .and. trim(word) /= 'learned') &
               call terminate("ladder must be 'linear' or 'learned'")
            gpu_learned = trim(word) == 'learned'
            ! optional training MCS and window count after the word
            if (gpu_learned) then
               read (value, *, iostat=ios) word, gpu_train_mcs, gpu_windows
               if (ios /= 0) then
                  gpu_windows = 4
                  read (value, *, iostat=ios) word, gpu_train_mcs
                  if (ios /= 0) gpu_train_mcs = 0
                  ios = 0
               end if
               if (gpu_train_mcs < 0 .or. gpu_windows < 1) &
                  call terminate('ladder: training MCS must be >= 0 and windows >= 1')
            end if
         case ('autotune')
            read (value, *, iostat=ios) gpu_autotune
         case ('launch')
            read (value, *, iostat=ios) gpu_spw, gpu_xcd
         case ('lanes')
            read (value, *, iostat=ios) gpu_lanes
         case ('graphs')
            read (value, *, iostat=ios) gpu_graphs
         case ('compat')
            read (value, *, iostat=ios) gpu_compat
            if (ios == 0 .and. (gpu_compat < 0 .or. gpu_compat > 15)) &
               call terminate('compat: a sum of the flags 1, 2, 4 and 8')
         case ('chunk')
            read (value, *, iostat=ios) gpu_chunk
            if (gpu_chunk < 0) call terminate('chunk must be >= 0')
         case default
            call terminate('Found unknown/missing information in gpu.ether: '//trim(adjustl(line)))
         end select
         if (ios /= 0) call terminate('cannot read the value of tag '//trim(key))
      end do
      close (u)
   end subroutine

   ! read_structure + frac_to_cart (src/read_structure.f90:51-106, src/frac_to_cart.f90:45-49)
   subroutine read_structure()
      character(len=400) :: line
      real(dp) :: wt, f(3)
      integer :: u, i, a, ios
      open (newunit=u, file='structure.vasp', status='old', action='read')
      read (u, '(a)') line
      read (u, *) wt
      do a = 1, 3
         read (u, *) lat(1:3, a)
      end do
      allocate (species_name(nspecies), ions(nspecies))
      read (u, *) (species_name(i), i=1, nspecies)
      do i = 1, nspecies
         species_name(i) = sym(species_name(i))
      end do
      read (u, *) (ions(i), i=1, nspecies)
      natoms = sum(ions)
      read (u, '(a)') line
      line = lower(adjustl(line))
      allocate (xyz(3, natoms))
      if (line(1:6) == 'direct') then
         do i = 1, natoms
            read (u, *, iostat=ios) f
            if (ios /= 0) call terminate('Incorrect structure file')
            xyz(:, i) = lat(:, 1)*f(1) + lat(:, 2)*f(2) + lat(:, 3)*f(3)
         end do
      else if (line(1:9) == 'cartesian') then
         do i = 1, natoms
            read (u, *, iostat=ios) xyz(:, i)
            if (ios /= 0) call terminate('Incorrect structure file')
         end do
      else
         call terminate('Incorrect structure file')
      end if
      close (u)
      if (.not. species_given) then
         allocate (species_incl(nspecies))
         species_incl = species_name
      end if
   end subroutine

end module ether_driver_input

program ether_gpu_driver
   use iso_c_binding
   use ether_gpu
   use ether_driver_input
   implicit none

   integer :: nb, n, i, j, k, l, s, a, b, q, nsh, npairs, u, ios, nscan, R, t, rep, coa_i
   integer :: start, nchunk, step, sw, rc, ovrr_iv, cnt, cap, first_counted, rd, crep, ncrep, compat_all
   logical :: chain
   integer, allocatable :: b_sp(:), b_idx(:), species(:), nn_count(:), nn_idx(:), gstart(:)
   integer(c_int), allocatable, target :: colour(:)
   integer(c_int) :: ncol_p
   integer(c_int), allocatable :: sp_partners(:)
   integer(c_int64_t) :: sp_count
   real(dp), allocatable :: sp_sums(:, :), sp_per(:, :, :), sp_mean(:, :), sp_sem(:, :), pos(:, :)
   character(len=20) :: hd(8)
   character(len=30) :: fname
   integer :: reclo, recls, ug, uo, us, uv
   integer(c_int), allocatable :: slot_of(:), rep_at(:)
   integer(c_int64_t), allocatable :: acc64(:), att64(:), trips64(:)
   integer, allocatable :: tions(:), inc(:)
   real(dp), allocatable :: sfin(:, :), phifin(:), ionrec(:, :), obs(:), mmv(:, :)
   real(dp), allocatable :: b_xyz(:, :), stagger(:), S0(:, :), phi0(:), shells(:), jx(:, :, :, :)
   real(dp), allocatable :: start_s(:, :), start_phi(:)
   real(dp), allocatable :: siav(:, :), temps(:), betas(:), mom(:, :), per(:, :, :), tnew(:)
   real(dp) :: lreport(8)
   real(dp), allocatable :: mean(:, :), sem(:, :), accp(:), stg_id(:)
   real(dp) :: kb, mub, gf, jv(3), ovrr_p, c1, r1
   real(dp) :: e, e2, e4, mm, m2, m4, nn
   character(len=8) :: s1, s2
   character(len=2) :: sa, sb
   character(len=400) :: line
   logical :: ok, ising
   type(c_ptr) :: ctx
   integer(c_int) :: swap_c, spw, xcdm
   integer(c_int), allocatable :: obc(:)
   integer(8) :: t0, t1, rate

   call read_species_count()
   call read_input()
   call read_gpu_input()
   call read_structure()
   ising = (model == 'ising')
   write (6, '(a)') ' ==> ETHER (MI355X engine) - inputs read'

   ! ---- basis of included species, in the order of the `species` tag
   !      (src/generate_supercell.f90:65-101)
   allocate (b_sp(natoms), b_idx(natoms), b_xyz(3, natoms))
   nb = 0
   do i = 1, size(species_incl)
      q = 0
      do j = 1, nspecies
         if (species_incl(i) == species_name(j)) then
            do k = q + 1, q + ions(j)
               nb = nb + 1
               b_sp(nb) = j
               b_idx(nb) = k
               b_xyz(:, nb) = xyz(:, k)
            end do
         end if
         q = q + ions(j)
      end do
   end do
   if (nb == 0) call terminate("Selected ions from 'in.ether' are not present in 'structure.vasp' file")

   ! ---- supercell: k (z) outer, j, i, basis l inner (src/generate_supercell.f90:115-132)
   n = sc(1)*sc(2)*sc(3)*nb
   allocate (species(n), stagger(n), S0(3, n), phi0(n), stg_id(nspecies), pos(3, n))
   stg_id = 1.0_dp
   inquire (file='staggered', exist=ok)
   if (stg .and. ok .and. .not. cb) then
      open (newunit=u, file='staggered', status='old', action='read')
      do
         read (u, *, iostat=ios) c1, sa
         if (ios /= 0) exit
         do q = 1, nspecies
            if (sym(sa) == species_name(q)) stg_id(q) = c1
         end do
      end do
      close (u)
   end if
   q = 0
   do k = 1, sc(3)
      do j = 1, sc(2)
         do i = 1, sc(1)
            do l = 1, nb
               q = q + 1
               species(q) = b_sp(l)
               pos(:, q) = lat(:, 1)*(i - 1) + lat(:, 2)*(j - 1) + lat(:, 3)*(k - 1) + b_xyz(:, l)
               if (cb) then
                  stagger(q) = merge(1.0_dp, -1.0_dp, mod(i + j + k + l, 2) == 0)
               else
                  stagger(q) = stg_id(b_sp(l))
               end if
            end do
         end do
      end do
   end do
   ! start state: Ising (-1)^i by file index; XYZ random per basis site
   ! (src/generate_supercell.f90:34-63), a function of SEED as the reference's
   ! random_seed(put=seed+rank) makes it (src/ether.f90:63), drawn by the
   ! library so that this driver and the Python one start alike; copied to
   ! every cell
   allocate (start_s(3, nb), start_phi(nb))
   rc = ether_start_state(int(seed_value, c_int64_t), merge(ETHER_MODEL_ISING, ETHER_MODEL_XYZ, ising), &
                          merge(1, 0, angle .and. .not. ising), nb, b_idx, start_s, start_phi)
   if (rc /= ETHER_OK) call terminate(ether_message(c_null_ptr))
   do l = 1, nb
      do q = l, n, nb
         S0(:, q) = start_s(:, l)
         phi0(q) = start_phi(l)
      end do
   end do
   deallocate (start_s, start_phi)

   ! ---- j_exchange (src/j_values.f90:39-87): J * s_a * s_b, symmetric
   inquire (file='j_exchange', exist=ok)
   if (.not. ok) call terminate("Required file 'j_exchange' is not present")
   open (newunit=u, file='j_exchange', status='old', action='read')
   read (u, *) nsh
   allocate (shells(nsh), jx(3, nspecies, nspecies, nsh))
   jx = 0.0_dp
   do s = 1, nsh
      read (u, *) shells(s), npairs
      do q = 1, npairs
         read (u, *) sa, sb, jv
         a = 0
         b = 0
         do k = 1, nspecies
            if (sym(sa) == species_name(k)) a = k
            if (sym(sb) == species_name(k)) b = k
         end do
         if (a == 0 .or. b == 0) call terminate('j_exchange names a species absent from structure.vasp')
         jx(:, b, a, s) = jv*spin(a)*spin(b)
         jx(:, a, b, s) = jx(:, b, a, s)
      end do
   end do
   close (u)

   ! ---- single_ion_anisotropy (src/get_sia_values.f90:32-62): D * s^2, XYZ only
   allocate (siav(3, nspecies))
   siav = 0.0_dp
   if (sia .and. .not. ising) then
      inquire (file='single_ion_anisotropy', exist=ok)
      if (.not. ok) call terminate("SIA is .TRUE. but file 'single_ion_anisotropy' is not found")
      open (newunit=u, file='single_ion_anisotropy', status='old', action='read')
      do
         read (u, *, iostat=ios) sa, jv
         if (ios /= 0) exit
         do k = 1, nspecies
            if (sym(sa) == species_name(k)) siav(:, k) = jv*spin(k)**2
         end do
      end do
      close (u)
   end if

   ! ---- parameters (src/parameters.f90:24-58)
   if (para) then
      jx = jx/j_para
      siav = siav/j_para
      kb = 1.0_dp
      mub = 1.0_dp
      gf = 1.0_dp
      if (zeeman) h = h/j_para
   else
      jx = jx/1000.0_dp
      siav = siav/1000.0_dp
      kb = 8.617333262d-5
      mub = 5.7883818060d-5
      gf = g_factor
   end if

   ! ---- neighbours (src/getting_nbd.f90:46-82 semantics)
   cap = nbd_capacity
   allocate (nn_count(n*nsh), nn_idx(n*nsh*cap), obc(3))
   do a = 1, 3
      obc(a) = merge(1, 0, bc(a) == 'o')
   end do
   ! on the GPU, with the periodic colouring; the host builder and colouring
   ! (the same lists and colours) where no GPU is visible
   allocate (colour(n))
   ncol_p = 0
   rc = ether_build_neighbours_device(lat, nb, b_xyz(:, 1:nb), sc, obc, nsh, shells, nbdfc, cap, &
                                      nn_count, nn_idx, colour, ncol_p)
   if (rc == ETHER_ERR_NODEV) then
      rc = ether_build_neighbours(lat, nb, b_xyz(:, 1:nb), sc, obc, nsh, shells, nbdfc, cap, nn_count, nn_idx)
      if (rc == ETHER_OK) rc = ether_colour_lattice(lat, nb, b_xyz(:, 1:nb), sc, obc, nsh, shells, nbdfc, &
                                                    colour, ncol_p)
   end if
   if (rc /= ETHER_OK) call terminate(ether_message(c_null_ptr))
   write (6, '(a)') ' ==> Neighbourhood informations are created now'

   ! ---- temperatures (src/ether.f90:79-84) and replicas r = rep*nscan + t
   nscan = nint((ht - lt)/tint) + 1
   allocate (temps(nscan))
   do t = 1, nscan
      temps(t) = ht - tint*(t - 1)
   end do
   R = nscan*nrepeat
   allocate (betas(R))
   do rep = 0, nrepeat - 1
      do t = 1, nscan
         betas(rep*nscan + t) = 1.0_dp/(kb*temps(t))
      end do
   end do

   ! ---- compat flags (gpu.ether `compat` and the ETHER_COMPAT variable); with
   !      ETHER_COMPAT_CHAIN_REPEATS the repeats run one after another on nscan
   !      replicas, each from fresh_spins of the previous repeat's final state
   !      (src/ether.f90:152-155), as the reference's rank does
   call get_environment_variable('ETHER_COMPAT', line, status=ios)
   q = 0
   if (ios == 0) read (line, *, iostat=ios) q
   compat_all = ior(q, gpu_compat)
   chain = iand(compat_all, ETHER_COMPAT_CHAIN_REPEATS) /= 0
   if (chain .and. gpu_learned) call terminate('ladder = learned does not run with chained repeats (compat 8)')
   rd = R
   if (chain) rd = nscan

   ! ---- context
   rc = ether_ctx_create(merge(ETHER_MODEL_ISING, ETHER_MODEL_XYZ, ising), n, rd, &
                         int(seed_value, c_int64_t), ctx)
   if (rc /= ETHER_OK) call terminate(ether_message(c_null_ptr))
   if (ncol_p > 0) then
      call chk(ether_set_lattice(ctx, nspecies, species, stagger, nsh, cap, nn_count, nn_idx, c_loc(colour)))
   else
      call chk(ether_set_lattice(ctx, nspecies, species, stagger, nsh, cap, nn_count, nn_idx, c_null_ptr))
   end if
   call chk(ether_set_exchange(ctx, jx, spin))
   call chk(ether_set_zeeman(ctx, merge(1, 0, zeeman), h, gf, mub))
   call chk(ether_set_sia(ctx, merge(1, 0, sia), siav))
   call chk(ether_set_proposal(ctx, merge(1, 0, angle), dphi))
   call chk(ether_set_compat(ctx, int(compat_all, c_int)))
   ! spiral-state partners (src/ether.f90:69, src/get_spiral_state.f90:47-82)
   sp_count = 0
   if (ssp .and. .not. ising) then
      allocate (sp_partners(n*spiral_capacity))
      rc = ether_build_spiral_partners(n, pos, ss_direc, ss_dis, ss_lat, spiral_capacity, &
                                       sp_partners, sp_count)
      if (rc /= ETHER_OK) call terminate(ether_message(c_null_ptr))
      call chk(ether_set_spiral(ctx, spiral_capacity, sp_partners))
   end if
   call chk(ether_commit(ctx))
   call chk(ether_set_graphs(ctx, merge(1, 0, gpu_graphs)))
   if (gpu_spw > 0) call chk(ether_set_launch(ctx, gpu_spw, gpu_xcd))
   if (gpu_lanes > 0) call chk(ether_set_lanes(ctx, gpu_lanes))
   call chk(ether_set_betas(ctx, betas(1:rd)))
   if (pt) then
      allocate (gstart(rd/nscan + 1))
      do rep = 0, rd/nscan
         gstart(rep + 1) = rep*nscan
      end do
      call chk(ether_set_ladders(ctx, rd/nscan, gstart))
   end if
   ! optional launch-shape autotuning: sweeps timed per candidate, from gpu.ether's
   ! `autotune` or the ETHER_AUTOTUNE environment variable
   call get_environment_variable('ETHER_AUTOTUNE', line, status=ios)
   q = gpu_autotune
   if (ios == 0) read (line, *, iostat=ios) q
   if (q > 0) then
      call chk(ether_autotune(ctx, int(q, c_int), spw, xcdm, c_null_ptr))
      write (6, '(a,i0,a,i0)') ' ==> autotuned sweep launch: sites per wave ', spw, ', XCD map ', xcdm
   end if
   ! generate_supercell state, then fresh_spins (src/ether.f90:155)
   call chk(ether_set_spins(ctx, -1, S0, phi0))
   call chk(ether_init_spins(ctx, ETHER_INIT_FRESH))
   call chk(ether_reset_moments(ctx))

   coa_i = nint(coa)
   if (abs(coa - coa_i) > 1.0d-12 .or. coa_i <= 0) call terminate('coa must be a positive integer')
   ! OVRR gate of src/ether.f90:163-167
   ovrr_iv = 0
   ovrr_p = 0.0_dp
   if (ovrr .and. .not. ising .and. .not. zeeman .and. .not. sia .and. ovrr_mcs > 0) then
      ovrr_iv = ovrr_mcs
      ovrr_p = ovrr_para
   end if
   write (6, '(a,i0,a,i0,a,i0,a)') ' ==> ', R, ' chains (', nscan, ' T x ', nrepeat, &
      ' repeats) resident on the GPU'

   ! ---- MCS loop in chunks; a STATUS file prints progress (src/ether.f90:180-186)
   call system_clock(t0, rate)
   swap_c = 0
   step = 1
   ! gpu.ether `ladder = learned`: the ladder is trained inside the equilibration
   ! MCS by the library (ether_ladder_train) and replaces the linear one for the
   ! rest of the run, where it beat the linear ladder in validation
   if (gpu_learned) then
      if (.not. pt) call terminate('ladder = learned needs parallel tempering (pt = .T.)')
      q = gpu_train_mcs
      if (q == 0) q = tmcs_eq/2
      if (q <= 0 .or. q + 3*(q/8) > tmcs_eq) &
         call terminate('ladder = learned: the training MCS (and three validation windows '// &
                        'of 1/8 of them) must lie within the equilibration MCS')
      allocate (tnew(nscan))
      call chk(ether_ladder_train(ctx, nscan, temps, kb, 1, q, gpu_windows, q/8, coa_i, ovrr_iv, &
                                  ovrr_p, exchange_interval, 1, 0, 1992_c_int64_t, swap_c, tnew, &
                                  lreport))
      temps = tnew
      do rep = 0, nrepeat - 1
         do t = 1, nscan
            betas(rep*nscan + t) = 1.0_dp/(kb*temps(t))
         end do
      end do
      step = 1 + nint(lreport(7))
      call chk(ether_reset_moments(ctx))
      if (lreport(5) > 0.5_dp) then
         write (6, '(a)') ' ==> learned tempering ladder kept after validation'
      else
         write (6, '(a)') ' ==> learned tempering ladder did not beat the linear one; linear kept'
      end if
   end if
   first_counted = step
   nchunk = 1000
   if (gpu_chunk > 0) nchunk = gpu_chunk
   allocate (mom(ETHER_MOM_COUNT, R))
   if (sp_count > 0) allocate (sp_sums(3, R))
   ! all repeats at once (rd = R replicas), or one after another (rd = nscan)
   ncrep = 1
   if (chain) ncrep = nrepeat
   do crep = 1, ncrep
      if (crep > 1) then
         ! the next repeat: fresh_spins of the state the last one left, zeroes('avg')
         call chk(ether_init_spins(ctx, ETHER_INIT_FRESH))
         call chk(ether_reset_moments(ctx))
         step = 1
         swap_c = 0
      end if
      do while (step <= tmcs)
         start = step
         q = min(nchunk, tmcs - step + 1)
         call chk(ether_run(ctx, start, q, tmcs_eq, coa_i, ovrr_iv, ovrr_p, &
                            merge(exchange_interval, 0, pt), swap_c))
         step = step + q
         inquire (file='STATUS', exist=ok)
         if (ok) then
            call chk(ether_synchronize(ctx))
            write (6, '(a,i0,a,i0)') ' STATUS: MCS ', step - 1, ' of ', tmcs
            open (newunit=u, file='STATUS')
            close (u, status='delete')
         end if
      end do
      call chk(ether_read_moments(ctx, mom(:, (crep - 1)*rd + 1:crep*rd)))
      if (sp_count > 0) call chk(ether_read_spiral(ctx, sp_sums(:, (crep - 1)*rd + 1:crep*rd), c_null_ptr))
   end do
   call system_clock(t1)

   ! ---- observables per repeat and over repeats (src/get_eng.f90:43-56,
   !      src/get_mag.f90:44-55, src/observables_per_repeat.f90:29-69)
   allocate (per(6, nrepeat, nscan), mean(6, nscan), sem(6, nscan), accp(nscan))
   accp = 0.0_dp
   do rep = 0, nrepeat - 1
      do t = 1, nscan
         q = rep*nscan + t
         nn = mom(1, q)
         e = mom(2, q)/nn
         e2 = mom(3, q)/nn
         e4 = mom(4, q)/nn
         mm = mom(5, q)/nn
         m2 = mom(6, q)/nn
         m4 = mom(7, q)/nn
         per(1, rep + 1, t) = e
         per(2, rep + 1, t) = 1.0_dp - (1.0_dp/3.0_dp)*(e4/(e2**2))
         per(3, rep + 1, t) = (betas(q)**2)*(e2 - e**2)*n
         per(4, rep + 1, t) = mm
         per(5, rep + 1, t) = 1.0_dp - (1.0_dp/3.0_dp)*(m4/(m2**2))
         per(6, rep + 1, t) = betas(q)*(m2 - mm**2)*n
         accp(t) = accp(t) + mom(8, q)
      end do
   end do
   r1 = 1.0_dp/(nrepeat*(nrepeat - 1.0_dp))
   do t = 1, nscan
      do k = 1, 6
         mean(k, t) = sum(per(k, :, t))/nrepeat
         sem(k, t) = sqrt(r1*sum((per(k, :, t) - mean(k, t))**2))
      end do
      accp(t) = accp(t)/(nrepeat*(tmcs - first_counted + 1))*100.0_dp
   end do

   ! spiral state per repeat and over repeats (src/get_spiral_state.f90:122-140,
   ! src/observables_per_repeat.f90:71-94)
   if (sp_count > 0) then
      allocate (sp_per(3, nrepeat, nscan), sp_mean(3, nscan), sp_sem(3, nscan))
      do rep = 0, nrepeat - 1
         do t = 1, nscan
            q = rep*nscan + t
            nn = mom(1, q)
            e = sp_sums(1, q)/nn
            e2 = sp_sums(2, q)/nn
            e4 = sp_sums(3, q)/nn
            sp_per(1, rep + 1, t) = e
            sp_per(2, rep + 1, t) = 1.5_dp - 0.5_dp*(e4/(e2**2))
            sp_per(3, rep + 1, t) = betas(q)*(e2 - e**2)*real(sp_count, dp)
         end do
      end do
      do t = 1, nscan
         do k = 1, 3
            sp_mean(k, t) = sum(sp_per(k, :, t))/nrepeat
            sp_sem(k, t) = sqrt(r1*sum((sp_per(k, :, t) - sp_mean(k, t))**2))
         end do
      end do
   end if

   ! output headers as src/write_output_files.f90:35-66 writes them
   open (newunit=u, file='magnetization.dat', status='replace', action='write')
   hd(1:7) = [character(len=20) :: '# 1 Temp.', '2 |M|', '3 χ', '4 Δ|M|', '5 Δχ', '6 U(|M|)', '7 ΔU(|M|)']
   write (u, 103) (hd(i), i = 1, 7)
   do t = 1, nscan
      write (u, 102) temps(t), mean(4, t), mean(6, t), sem(4, t), sem(6, t), mean(5, t), sem(5, t)
   end do
   close (u)
   open (newunit=u, file='energy.dat', status='replace', action='write')
   hd = [character(len=20) :: '# 1 Temp.', '2 E', '3 Cv/kB', '4 ΔE', '5 Δ(Cv/kB)', '6 U(E)', '7 ΔU(E)', &
         '8 Accpt. (%)']
   if (para) then
      hd(3) = '3 Cv'
      hd(5) = '5 ΔCv'
   end if
   write (u, 104) (hd(i), i = 1, 8)
   do t = 1, nscan
      write (u, 102) temps(t), mean(1, t), mean(3, t), sem(1, t), sem(3, t), mean(2, t), sem(2, t), accp(t)
   end do
   close (u)
   open (newunit=u, file='spiral_state_parameters.dat', status='replace', action='write')
   if (sp_count > 0) then
      hd(1:7) = [character(len=20) :: '# 1 Temp.', '2 |ψ|', '3 ψ', '4 Δ|ψ|', '5 Δψ', '6 U(|ψ|)', '7 ΔU(|ψ|)']
      write (u, 103) (hd(i), i = 1, 7)
      do t = 1, nscan
         write (u, 102) temps(t), sp_mean(1, t), sp_mean(3, t), sp_sem(1, t), sp_sem(3, t), &
            sp_mean(2, t), sp_sem(2, t)
      end do
   end if
   close (u)
102 format(f11.5, 1x, es12.5, 1x, es12.5, 1x, es12.5, 1x, es12.5, 1x, es12.5, 1x, es12.5, 1x, f7.2)
103 format(A13, A13, A13, 1x, A13, 1x, A13, 2x, A13, A13)
104 format(A13, A13, A13, A13, 1x, A13, 1x, A13, A13, A13)

   ! ---- the configuration at each temperature after the last repeat, and the
   !      files built from it: ETHER.obs / ETHER.spn direct-access records
   !      (src/process_observables.f90:34-77, src/ether.f90:216-231), gss.dat
   !      (src/write_gss.f90:34-78), moment_vectors.dat
   !      (src/write_output_files.f90:66-80) and <T>spK.xsf
   !      (src/write_spins_at_K.f90:33-66).  Replica r = rep*nscan + t holds slot
   !      slot_of(r); the one in slot rd - nscan + t (the last repeat's ladder) is at
   !      temperature t.
   allocate (slot_of(rd), rep_at(rd), acc64(rd), att64(rd), trips64(rd))
   allocate (sfin(3, n), phifin(n), ionrec(0:8, n), obs(20 + 3*nspecies), mmv(3, nspecies), tions(nspecies))
   call chk(ether_pt_state(ctx, slot_of, acc64, att64, trips64))
   do q = 1, rd
      rep_at(slot_of(q) + 1) = q - 1
   end do
   tions = 0
   do i = 1, n
      tions(species(i)) = tions(species(i)) + 1
   end do
   cnt = 0
   allocate (inc(size(species_incl)))
   do i = 1, size(species_incl)
      do j = 1, nspecies
         if (species_incl(i) == species_name(j)) then
            cnt = cnt + 1
            inc(cnt) = j
         end if
      end do
   end do
   inquire (iolength=reclo) obs
   inquire (iolength=recls) ionrec
   open (newunit=uo, file='ETHER.obs', form='unformatted', access='direct', recl=reclo, status='replace')
   open (newunit=us, file='ETHER.spn', form='unformatted', access='direct', recl=recls, status='replace')
   open (newunit=ug, file='gss.dat', status='replace', action='write')
   write (ug, '(I5,1x, I5, 1x, I5, 1x, I5)') nscan, sc
   write (ug, '(I5,1x, I5, 1x, I5, 1x, I5, 1x, I5, 1x, I5)') 1, 1, 1, sc
   write (ug, '(I5,1x, I5, f11.5, 1x, f11.5, 1x, f11.5,1x, A6)') nb, nspecies, ht, lt, tint
   fname = ' K'
   if (para) fname = ' KbT/J'
   write (ug, '(A20)') fname
   write (ug, '(I8,1x,I8,1x,I8,1x,I8,1x,I8)') tions
   do i = 1, 3
      write (ug, '(f10.5,1x,f10.5,1x,f10.5)') lat(:, i)
   end do
   write (ug, *) '# lattice points'
   do i = 1, n
      write (ug, '(f11.6, 1x, f11.6, 1x, f11.6, 1X, i8)') pos(:, i), i
   end do
   open (newunit=uv, file='moment_vectors.dat', status='replace', action='write')
   write (uv, *) '# Showing moment vectors (Mx, My, Mz) for species(sp):'
   write (uv, 105) (species_name(inc(i)), i = 1, cnt)
   write (uv, *) '# Temp, sp1(Mx,My,Mz), sp1(Mx,My,Mz),..spN(Mx,My,Mz), sp1|M|, sp2|M|,..spN|M|'
   do t = 1, nscan
      call chk(ether_get_spins(ctx, rep_at(rd - nscan + t), sfin, phifin))
      mmv = 0.0_dp
      do i = 1, n
         mmv(:, species(i)) = mmv(:, species(i)) + stagger(i)*sfin(:, i)*spin(species(i))
      end do
      obs = 0.0_dp
      obs(1:14) = [temps(t), mean(4, t), mean(6, t), sem(4, t), sem(6, t), mean(5, t), sem(5, t), &
                   mean(1, t), mean(3, t), sem(1, t), sem(3, t), mean(2, t), sem(2, t), accp(t)]
      if (sp_count > 0) obs(15:20) = [sp_mean(1, t), sp_mean(3, t), sp_sem(1, t), sp_sem(3, t), &
                                      sp_mean(2, t), sp_sem(2, t)]
      do j = 1, nspecies
         obs(20 + 3*(j - 1) + 1:20 + 3*j) = mmv(:, j)
      end do
      write (uo, rec=t) obs
      do i = 1, n
         ionrec(0, i) = real(i, dp)
         ionrec(1:3, i) = sfin(:, i)
         ionrec(4, i) = real(species(i), dp)
         ionrec(5, i) = phifin(i)
         ionrec(6:8, i) = pos(:, i)
      end do
      write (us, rec=t) ionrec
      write (ug, '(f11.5)') temps(t)
      do i = 1, n
         write (ug, '(f11.6, 1x, f11.6, 1x, f11.6, 1X, i8)') sfin(:, i), species(i)
      end do
      write (uv, 101) temps(t), ((mmv(:, inc(i))/tions(inc(i))), i = 1, cnt), &
         (sqrt(dot_product(mmv(:, inc(i)), mmv(:, inc(i))))/tions(inc(i)), i = 1, cnt)
      write (fname, '(f9.4,"spK.xsf")') temps(t)
      open (newunit=u, file=adjustl(fname), status='replace', action='write')
      write (u, *) 'CRYSTAL'
      write (u, *) 'PRIMVEC'
      do a = 1, 3
         write (u, *) lat(:, a)*sc(a)
      end do
      write (u, *) 'CONVEC'
      do a = 1, 3
         write (u, *) lat(:, a)*sc(a)
      end do
      write (u, *) 'PRIMCOORD'
      write (u, *) n, " 1"
      do i = 1, n
         write (u, 106) species_name(species(i)), pos(:, i), sfin(:, i)
      end do
      close (u)
   end do
   close (uo)
   close (us)
   close (ug)
   close (uv)
101 format(f11.5, 2X, 500f9.3)
105 format(' #	', 5A3)
106 format(A5, 1x, f13.7, 1x, f13.7, 1x, f13.7, 1x, f13.7, 1x, f13.7, 1x, f13.7)

   write (s1, '(f8.3)') real(t1 - t0, dp)/real(rate, dp)
   write (s2, '(es8.2)') real(R, dp)*n*tmcs/(real(t1 - t0, dp)/real(rate, dp))
   write (6, '(a)') ' ==> MCS loop: '//trim(adjustl(s1))//' s, '//trim(adjustl(s2))//' spin updates/s'
   write (6, '(a)') ' ==> energy.dat, magnetization.dat, spiral_state_parameters.dat, ETHER.obs, '// &
      'ETHER.spn, gss.dat, moment_vectors.dat and spK.xsf files written'
   call ether_ctx_destroy(ctx)

contains

   subroutine chk(code)
      integer(c_int), intent(in) :: code
      if (code /= ETHER_OK) call terminate(ether_message(ctx))
   end subroutine

end program ether_gpu_driver
