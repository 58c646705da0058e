! ocfd_dropin.f90 -- the reference solver's hot-path routines served by
! libocfd_hip.so, for the reference driver NS_solver (OCFD_NS_Solver.f90:5-90)
! left unchanged.
!
! The executable is linked from the reference's own objects with
! `ld --wrap=<symbol>` for every routine below: a call from NS_solver to
! du_inviscous_ lands in __wrap_du_inviscous_ here, and __real_<symbol> is the
! reference's original.  The state lives on the GPU from the first OCFD_bc
! (end of initialisation, OCFD_NS_Solver.f90:15) on; host arrays are refreshed
! from the device only where host code reads them (show_flow_msg,
! OCFD_analysis); save_flow_data is written by the library from the device.
!
!   exchange_boundary_xyz(d..T)  -> ocfd_exchange_primitives (OCFD_mpi.f90:183-196)
!   comput_Rhybrid               -> ocfd_comput_rhybrid      (OCFD_SetHybrid.f90:2-152)
!   du_inviscous                 -> ocfd_du_inviscous        (OCFD_inviscous.f90:53-203)
!   du_inviscous_Character       -> ocfd_du_inviscous_character (OCFD_inviscous_character.f90:3-159)
!   du_viscous                   -> ocfd_du_viscous          (OCFD_viscous.f90:29-131)
!   OCFD_time_adv_RK3(KRK)       -> ocfd_save_fn (KRK=1: the fn=f loop of
!                                   OCFD_NS_Solver.f90:23-31) + ocfd_time_adv_rk3
!   comput_duvwT                 -> ocfd_comput_duvwt        (OCFD_Time_Adv.f90:40-63)
!   OCFD_bc                      -> ocfd_bc                  (OCFD_BC.f90:2-99); with
!                                   BoundaryCondition = BC_User_Def the
!                                   user's own ocfd_bc_User runs on the host
!                                   (OCFD_BC.f90:18-19): the state is copied
!                                   down, the reference's OCFD_bc runs, the
!                                   state is copied back.  OCFD_HIP_USER_BC=device
!                                   instead runs the library's built-in hook on
!                                   the device (isothermal or adiabatic no-slip
!                                   wall at j=1, free stream at j=ny); only a
!                                   case whose ocfd_bc_User is that wall may ask
!                                   for it.
!   filtering(f)                 -> ocfd_advance_clock + ocfd_filtering (OCFD_filtering.f90:4-70)
!   show_flow_msg                -> ocfd_show_flow_msg (OCFD_Time_Adv.f90:65-104), printed
!                                   and logged in the reference's formats
!   OCFD_analysis                -> ocfd_ana_box when every analysis due is ana_box
!                                   (OCFD_ANA.f90:178-311); any other on the host
!
! The stage's routines run in stage fusion (ocfd_set_stage_fusion): the
! library merges du_viscous, OCFD_time_adv_RK3 and comput_duvwT into the
! kernels of its whole-step path, with the same results.  Environment:
!   OCFD_HIP_STAGE_FUSION=0  one kernel sequence per routine instead
!   OCFD_HIP_DIAG=host       show_flow_msg / OCFD_analysis on the host after a
!                            download (their printed sums then equal the
!                            reference's to the last digit; the device sums
!                            add in another order, within 1e-13)
module ocfd_dropin_state
   use iso_c_binding
   implicit none
   type(c_ptr), save :: hdev = c_null_ptr
   ! .true. while host code that reads the state runs (its own halo
   ! exchanges must then act on the host arrays)
   logical, save :: host_section = .false.
   ! BC_User_Def served by the library's device hook (OCFD_HIP_USER_BC=device)
   ! instead of the user's host routine
   logical, save :: user_bc_on_device = .false.
   ! show_flow_msg / ana_box from the device sums (OCFD_HIP_DIAG /= host)
   logical, save :: diag_on_device = .true.
end module ocfd_dropin_state

! Creates the device block from the reference's parameters and uploads the
! metrics and the initial state (after init3d and the first OCFD_bc).
subroutine ocfd_dropin_attach()
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   type(ocfd_config) :: cfg
   type(c_ptr) :: mets(19)
   integer :: dev, fusion
   character(len=16) :: env
   integer :: st

   if (np_size /= 1) then
      print *, 'libocfd_hip drop-in: one MPI rank per block only (np_size =', np_size, ')'
      stop
   end if
   cfg%nx = nx_global; cfg%ny = ny_global; cfg%nz = nz_global; cfg%lap = LAP
   cfg%iperiodic = [Para%Iperiodic_X, Para%Iperiodic_Y, Para%Iperiodic_Z]
   cfg%scheme_invis = Scheme%Scheme_Invis
   cfg%scheme_vis = Scheme%Scheme_Vis
   cfg%flux_splitting = Para%Flux_Splitting
   cfg%if_scheme_character = Para%IF_Scheme_Character
   cfg%if_viscous = Para%IF_Viscous
   cfg%ibc = Para%IBC
   cfg%if_mass_force = Para%IF_Mass_Force
   cfg%scheme_boundary = Scheme%Scheme_boundary
   cfg%ghost_cell = Para%Ghost_cell
   cfg%nfiltering = Para%Nfiltering
   cfg%istep_show = Para%Istep_Show
   cfg%precision = OCFD_PREC_FP64
   cfg%metric_mode = OCFD_METRIC_FIELD
   cfg%reserved_i = 0
   cfg%ma = Para%Ma; cfg%re = Para%Re; cfg%gamma = Para%gamma; cfg%pr = Para%Pr
   cfg%ref_amu_t0 = Para%Ref_Amu_T0; cfg%dt = Para%dt
   cfg%ud7l_diss = Scheme%UD7L_Diss
   cfg%mass_force = Para%Mass_Force
   cfg%bc_para = Para%BC_para
   cfg%hybrid_para = Scheme%Hybrid_para
   cfg%filter_para = Para%Filter
   cfg%reserved_d = 0.d0

   dev = 0
   call get_environment_variable('OCFD_HIP_DEVICE', env, status=st)
   if (st == 0) read (env, *) dev
   user_bc_on_device = .false.
   call get_environment_variable('OCFD_HIP_USER_BC', env, status=st)
   if (st == 0) then
      if (trim(env) == 'device') then
         user_bc_on_device = .true.
      else if (trim(env) /= 'host') then
         print *, 'libocfd_hip drop-in: OCFD_HIP_USER_BC must be host or device, not ', trim(env)
         stop
      end if
   end if
   diag_on_device = .true.
   call get_environment_variable('OCFD_HIP_DIAG', env, status=st)
   if (st == 0) then
      if (trim(env) == 'host') then
         diag_on_device = .false.
      else if (trim(env) /= 'device') then
         print *, 'libocfd_hip drop-in: OCFD_HIP_DIAG must be host or device, not ', trim(env)
         stop
      end if
   end if
   fusion = 1
   call get_environment_variable('OCFD_HIP_STAGE_FUSION', env, status=st)
   if (st == 0) read (env, *) fusion
   call ocfd_check(ocfd_create(cfg, int(dev, c_int), hdev), 'ocfd_create')
   call ocfd_check(ocfd_set_stage_fusion(hdev, int(fusion, c_int32_t)), 'ocfd_set_stage_fusion')

   mets = [ocfd_addr(Akx), ocfd_addr(Aky), ocfd_addr(Akz), ocfd_addr(Aix), ocfd_addr(Aiy), &
           ocfd_addr(Aiz), ocfd_addr(Asx), ocfd_addr(Asy), ocfd_addr(Asz), ocfd_addr(Ajac), &
           ocfd_addr(Akx1), ocfd_addr(Aky1), ocfd_addr(Akz1), ocfd_addr(Aix1), ocfd_addr(Aiy1), &
           ocfd_addr(Aiz1), ocfd_addr(Asx1), ocfd_addr(Asy1), ocfd_addr(Asz1)]
   call ocfd_check(ocfd_upload_metrics(hdev, mets), 'ocfd_upload_metrics')
   call ocfd_check(ocfd_upload_primitives(hdev, d, u, v, w, T, int(Istep, c_int32_t), tt), &
                   'ocfd_upload_primitives')
   call ocfd_check(ocfd_upload_f(hdev, f), 'ocfd_upload_f')
   if (Para%IBC == BC_BoundaryLayer .or. Para%IBC == BC_SweptCorner) call ocfd_dropin_bc_data()
   if (my_id .eq. 0) print *, 'libocfd_hip: block on GPU', dev
   if (Para%IBC == BC_User_Def .and. my_id .eq. 0) then
      if (user_bc_on_device) then
         print *, 'libocfd_hip: BC_User_Def by the library''s device wall hook'
      else
         print *, 'libocfd_hip: BC_User_Def by the host routine ocfd_bc_User'
      end if
   end if
end subroutine ocfd_dropin_attach

! The data of ocfd_bc_BoundaryLayer / ocfd_bc_swept_corner that the first
! (host) OCFD_bc prepared in module BC_data (OCFD_BC_Boundarylayers.f90:1-15,
! 171-397), handed to the library; 1-d inlet/upper data repeated over k.
subroutine ocfd_dropin_bc_data()
   use flow_data
   use BC_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   real(c_double), allocatable, target :: inl(:, :, :), upp(:, :, :), pert(:, :)
   integer(c_int32_t), allocatable, target :: nonref(:, :), wmask(:, :)
   type(c_ptr) :: p_inl, p_upp, p_wall
   integer :: inlet_mode, upper_mode, j, k, i, m
   real(c_double), parameter :: epsl = 1.d-7

   if (Para%IBC == BC_BoundaryLayer) then
      inlet_mode = nint(Para%BC_para(1))
      upper_mode = nint(Para%BC_para(2))
   else
      inlet_mode = 1
      upper_mode = 0
   end if
   p_inl = c_null_ptr
   p_upp = c_null_ptr
   p_wall = c_null_ptr
   if (inlet_mode == 1 .or. inlet_mode == 2) then
      allocate (inl(ny, nz, 5))
      do m = 1, 5
      do k = 1, nz
      do j = 1, ny
         if (inlet_mode == 1) then
            inl(j, k, m) = BC%flow_inlet1d(j, m)
         else
            inl(j, k, m) = BC%flow_inlet2d(j, k, m)
         end if
      end do
      end do
      end do
      p_inl = c_loc(inl)
   end if
   if (upper_mode == 1 .or. upper_mode == 2) then
      allocate (upp(nx, nz, 5))
      do m = 1, 5
      do k = 1, nz
      do i = 1, nx
         if (upper_mode == 1) then
            upp(i, k, m) = BC%flow_upper1d(i, m)
         else
            upp(i, k, m) = BC%flow_upper2d(i, k, m)
         end if
      end do
      end do
      end do
      p_upp = c_loc(upp)
   end if
   allocate (nonref(nx, nz), pert(nx, nz))
   nonref = BC%NonReflect_upper
   pert = BC%wall_pertb
   if (Para%IBC == BC_SweptCorner) then
      ! no-slip wall points of ocfd_bc_swept_corner (OCFD_BC_Boundarylayers.f90:517)
      allocate (wmask(nx, nz))
      do k = 1, nz
      do i = 1, nx
         wmask(i, k) = 0
         if ((Azz(i, 1, k) >= -epsl .and. Azz(i, 1, k) <= Para%Bc_para(3) + epsl) .or. Ayy(i, 1, k) <= epsl) &
            wmask(i, k) = 1
      end do
      end do
      p_wall = c_loc(wmask)
   end if
   call ocfd_check(ocfd_set_bc_boundarylayer(hdev, Para%AoA, p_inl, p_upp, c_loc(nonref), c_loc(pert), p_wall), &
                   'ocfd_set_bc_boundarylayer')
end subroutine ocfd_dropin_bc_data

! Device state -> host arrays d,u,v,w,T (halo'd) and f.
subroutine ocfd_dropin_download()
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   call ocfd_check(ocfd_download_primitives(hdev, d, u, v, w, T), 'ocfd_download_primitives')
   call ocfd_check(ocfd_download_f(hdev, f), 'ocfd_download_f')
end subroutine ocfd_dropin_download

! Host arrays d,u,v,w,T (halo'd) and f -> device state.
subroutine ocfd_dropin_upload()
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   call ocfd_check(ocfd_upload_primitives(hdev, d, u, v, w, T, int(Istep, c_int32_t), tt), &
                   'ocfd_upload_primitives')
   call ocfd_check(ocfd_upload_f(hdev, f), 'ocfd_upload_f')
end subroutine ocfd_dropin_upload

subroutine wrap_ocfd_bc() bind(C, name="__wrap_ocfd_bc_")
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   interface
      subroutine real_ocfd_bc() bind(C, name="__real_ocfd_bc_")
      end subroutine
   end interface
   if (.not. c_associated(hdev)) then
      ! initial boundary condition on the host, then the state moves to the GPU
      call real_ocfd_bc()
      call ocfd_dropin_attach()
   else if (host_section) then
      call real_ocfd_bc()
   else if (Para%IBC == BC_User_Def .and. .not. user_bc_on_device) then
      ! the user's routine reads and writes host arrays: the whole OCFD_bc
      ! (hook, boundary-plane f, ghost cells) runs on a copy of the state
      call ocfd_dropin_download()
      host_section = .true.
      call real_ocfd_bc()
      host_section = .false.
      call ocfd_dropin_upload()
   else
      if (Para%IBC == BC_BoundaryLayer .or. Para%IBC == BC_SweptCorner) call ocfd_dropin_bc_time()
      call ocfd_check(ocfd_bc(hdev), 'ocfd_bc')
   end if
end subroutine wrap_ocfd_bc

! The wall perturbation's time factor at this call, from the reference's own
! get_ht_multifrequancy (its random phases are drawn on the first, host, call)
subroutine ocfd_dropin_bc_time()
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   real(kind=OCFD_REAL_KIND) :: ht
   integer :: mt
   real(kind=OCFD_REAL_KIND) :: beta
   if (Para%IBC == BC_BoundaryLayer) then
      mt = nint(Para%Bc_para(11))
      beta = Para%Bc_para(14)
   else
      mt = nint(Para%Bc_para(8))
      beta = Para%Bc_para(10)
   end if
   call get_ht_multifrequancy(ht, tt, mt, beta)
   call ocfd_check(ocfd_set_bc_time_factor(hdev, real(ht, c_double)), 'ocfd_set_bc_time_factor')
end subroutine ocfd_dropin_bc_time

subroutine wrap_exchange_boundary_xyz(a) bind(C, name="__wrap_exchange_boundary_xyz_")
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   real(c_double), target :: a(*)
   type(c_ptr) :: pa
   interface
      subroutine real_exchange_boundary_xyz(a) bind(C, name="__real_exchange_boundary_xyz_")
         import :: c_double
         real(c_double) :: a(*)
      end subroutine
   end interface
   if (.not. c_associated(hdev) .or. host_section) then
      call real_exchange_boundary_xyz(a)
      return
   end if
   pa = c_loc(a(1))
   if (c_associated(pa, ocfd_addr(d))) then
      ! NS_solver exchanges d, u, v, w, T in this order: all five on the first
      call ocfd_check(ocfd_exchange_primitives(hdev), 'ocfd_exchange_primitives')
   else if (c_associated(pa, ocfd_addr(u)) .or. c_associated(pa, ocfd_addr(v)) .or. &
            c_associated(pa, ocfd_addr(w)) .or. c_associated(pa, ocfd_addr(T))) then
      continue
   else
      call real_exchange_boundary_xyz(a)
   end if
end subroutine wrap_exchange_boundary_xyz

subroutine wrap_comput_rhybrid() bind(C, name="__wrap_comput_rhybrid_")
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   integer(c_int64_t), target :: counts(4)
   if (mod(Istep, Para%Istep_show) .eq. 0) then
      call ocfd_check(ocfd_comput_rhybrid(hdev, c_loc(counts)), 'ocfd_comput_rhybrid')
      ! the message of OCFD_SetHybrid.f90:134
      if (my_id .eq. 0) print *, "---the percent of 3 schemes (linear ,WENO7,WENO5 are: ", &
         counts(1:3)/(1.d0*counts(4))
   else
      call ocfd_check(ocfd_comput_rhybrid(hdev, c_null_ptr), 'ocfd_comput_rhybrid')
   end if
end subroutine wrap_comput_rhybrid

subroutine wrap_du_inviscous() bind(C, name="__wrap_du_inviscous_")
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   call ocfd_check(ocfd_du_inviscous(hdev), 'ocfd_du_inviscous')
end subroutine wrap_du_inviscous

subroutine wrap_du_inviscous_character() bind(C, name="__wrap_du_inviscous_character_")
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   call ocfd_check(ocfd_du_inviscous_character(hdev), 'ocfd_du_inviscous_character')
end subroutine wrap_du_inviscous_character

subroutine wrap_du_viscous() bind(C, name="__wrap_du_viscous_")
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   call ocfd_check(ocfd_du_viscous(hdev), 'ocfd_du_viscous')
end subroutine wrap_du_viscous

subroutine wrap_ocfd_time_adv_rk3(krk) bind(C, name="__wrap_ocfd_time_adv_rk3_")
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   integer(c_int32_t) :: krk
   ! fn = f of the step start (OCFD_NS_Solver.f90:23-31) happens on the device
   if (krk == 1) call ocfd_check(ocfd_save_fn(hdev), 'ocfd_save_fn')
   call ocfd_check(ocfd_time_adv_rk3(hdev, krk), 'ocfd_time_adv_rk3')
end subroutine wrap_ocfd_time_adv_rk3

subroutine wrap_comput_duvwt() bind(C, name="__wrap_comput_duvwt_")
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   integer(c_int) :: nneg
   nneg = ocfd_comput_duvwt(hdev)
   call ocfd_check(nneg, 'ocfd_comput_duvwt')
   if (nneg > 0) then
      ! handle_NegativeT (OCFD_handle_NegativeT.f90:1-17): cells repaired on
      ! the device; the run stops beyond 10 of them
      print *, "Negative T found !", nneg, " cells (repaired)"
      if (nneg > 10) then
         print *, "Number of Negative Temperature Points >  Limit, STOP !"
         stop
      end if
   end if
end subroutine wrap_comput_duvwt

subroutine wrap_filtering(fa) bind(C, name="__wrap_filtering_")
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   real(c_double) :: fa(*)
   integer(c_int32_t) :: istep_dev
   real(c_double) :: tt_dev
   ! NS_solver advances Istep and tt right before filtering(f)
   ! (OCFD_NS_Solver.f90:64-67); the device clock follows here
   call ocfd_check(ocfd_advance_clock(hdev), 'ocfd_advance_clock')
   call ocfd_check(ocfd_get_time(hdev, istep_dev, tt_dev), 'ocfd_get_time')
   if (istep_dev /= Istep) then
      print *, 'libocfd_hip drop-in: device step', istep_dev, ' differs from Istep', Istep
      stop
   end if
   call ocfd_check(ocfd_filtering(hdev), 'ocfd_filtering')
end subroutine wrap_filtering

subroutine wrap_show_flow_msg(wall_time) bind(C, name="__wrap_show_flow_msg_")
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   real(c_double) :: wall_time
   real(c_double) :: e(3), t0
   interface
      subroutine real_show_flow_msg(wall_time) bind(C, name="__real_show_flow_msg_")
         import :: c_double
         real(c_double) :: wall_time
      end subroutine
   end interface
   if (.not. diag_on_device) then
      call ocfd_dropin_download()
      host_section = .true.
      call real_show_flow_msg(wall_time)
      host_section = .false.
      return
   end if
   ! E, K, S averaged over the block (device reduction), reported as
   ! show_flow_msg reports them (OCFD_Time_Adv.f90:91-101)
   t0 = wall_time
   wall_time = MPI_wtime()
   call ocfd_check(ocfd_show_flow_msg(hdev, e), 'ocfd_show_flow_msg')
   if (my_id .eq. 0) then
      print *, 'Istep=', Istep, 'tt=', tt
      print *, 'CPU time per', Para%Istep_show, 'step is:    ', wall_time - t0
      write (*, "(3A25)") "Averaged Total-energy E", "Kinetic-energy K", "Entropy S"
      write (*, "(3E25.15)") e(1), e(2), e(3)
      open (66, file='opencfd.log', position='append')
      write (66, *) 'Istep=', Istep, 'tt=', tt, 'CPU time is', wall_time - t0
      write (66, "(3A25)") "Averaged Total-energy E", "Kinetic-energy K", "Entropy S"
      write (66, "(3E25.15)") e(1), e(2), e(3)
      close (66)
   end if
end subroutine wrap_show_flow_msg

! ana_box's report from the device statistics (OCFD_ANA.f90:295-310):
! s(1:20) in the column order of statistics.dat
subroutine ocfd_dropin_ana_box()
   use flow_data, only: tt, my_id  ! (its constant OCFD_ANA_BOX would hide the entry point)
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   real(c_double) :: s(20)
   call ocfd_check(ocfd_ana_box(hdev, s), 'ocfd_ana_box')
   s(1) = tt
   if (my_id .eq. 0) then
      write (*, '(4(1X,A13))') "tt", "Re_lamda", "Re_lamdax", "Amt"
      write (*, '(4(1X,E13.6E2))') s(1:4)
      write (*, '(4(1X,A13))') "vv_rms", "u_rms", "v_rms", "w_rms"
      write (*, '(4(1X,E13.6E2))') s(5:8)
      write (*, '(4(1X,A13))') "u_skewness", "u_flatness", "ux_skewness", "ux_flatness"
      write (*, '(4(1X,E13.6E2))') s(9:12)
      write (*, '(4(1X,A13))') "E_kinetic", "Enstrophy", "d_rms", "T_aver"
      write (*, '(4(1X,E13.6E2))') s(13:16)
      write (*, '(4(1X,A13))') "Amu_aver", "d_aver", "Ux_rms", "Alamdax"
      write (*, '(4(1X,E13.6E2))') s(17:20)
      open (33, file='statistics.dat', position='append')
      write (33, '(20(E20.13E2))') s
      close (33)
   end if
end subroutine ocfd_dropin_ana_box

subroutine wrap_ocfd_analysis() bind(C, name="__wrap_ocfd_analysis_")
   use flow_data
   use ocfd_dropin_state
   implicit none
   integer :: m, kstep
   logical :: due, box_only
   interface
      subroutine real_ocfd_analysis() bind(C, name="__real_ocfd_analysis_")
      end subroutine
   end interface
   ! the schedule test of OCFD_analysis (OCFD_ANA.f90:9-14)
   due = .false.
   box_only = .true.
   do m = 1, Para%ANA_Number
      kstep = nint(Para%ANA_Para(2, m))
      if (mod(Istep, kstep) .eq. 0) then
         due = .true.
         if (nint(Para%ANA_Para(1, m)) /= OCFD_ANA_BOX) box_only = .false.
      end if
   end do
   if (.not. due) return
   if (diag_on_device .and. box_only) then
      ! the due analyses, in their order, all ana_box
      do m = 1, Para%ANA_Number
         kstep = nint(Para%ANA_Para(2, m))
         if (mod(Istep, kstep) .eq. 0) call ocfd_dropin_ana_box()
      end do
      return
   end if
   call ocfd_dropin_download()
   host_section = .true.
   call real_ocfd_analysis()
   host_section = .false.
end subroutine wrap_ocfd_analysis

! save_flow_data (OCFD_IO.f90:44-64): the library writes the snapshot from
! the device state in the reference's record layout, with the reference's
! file name and messages
subroutine wrap_save_flow_data() bind(C, name="__wrap_save_flow_data_")
   use flow_data
   use ocfd_hip
   use ocfd_dropin_state
   implicit none
   interface
      subroutine real_save_flow_data() bind(C, name="__real_save_flow_data_")
      end subroutine
   end interface
   character(len=100) :: filename1
   if (.not. c_associated(hdev)) then
      call real_save_flow_data()
      return
   end if
   write (filename1, "('OCFD'I8.8'.dat')") Istep
   if (my_id .eq. 0) print *, 'write data file:', filename1
   call ocfd_check(ocfd_save_flow_data(hdev, trim(filename1)//c_null_char, 0_c_int32_t), 'ocfd_save_flow_data')
   if (my_id .eq. 0) print *, 'write data ok'
end subroutine wrap_save_flow_data
