! ocfd_hip_mod.f90 -- Fortran 2003 (iso_c_binding) interface to the C ABI of
! libocfd_hip.so declared in include/ocfd_hip.h.  Every C entry point returns
! an int status: 0 = OK, < 0 = configuration / HIP error (text from
! ocfd_last_error), > 0 = numerical event (count of T <= 0 cells).
module ocfd_hip
   use iso_c_binding
   implicit none

   integer(c_int32_t), parameter :: OCFD_PREC_FP64 = 0, OCFD_PREC_MIXED = 1
   integer(c_int32_t), parameter :: OCFD_METRIC_FIELD = 0, OCFD_METRIC_UNIFORM = 1

   ! struct ocfd_config (same member order, sizes and alignment)
   type, bind(C) :: ocfd_config
      integer(c_int32_t) :: nx, ny, nz, lap
      integer(c_int32_t) :: iperiodic(3)
      integer(c_int32_t) :: scheme_invis, scheme_vis, flux_splitting, if_scheme_character
      integer(c_int32_t) :: if_viscous, ibc, if_mass_force
      integer(c_int32_t) :: scheme_boundary(6), ghost_cell(6)
      integer(c_int32_t) :: nfiltering, istep_show, precision, metric_mode
      integer(c_int32_t) :: reserved_i(8)
      real(c_double) :: ma, re, gamma, pr, ref_amu_t0, dt
      real(c_double) :: ud7l_diss
      real(c_double) :: mass_force(3)
      real(c_double) :: bc_para(100)
      real(c_double) :: hybrid_para(100)
      real(c_double) :: filter_para(100, 10)   ! Filter(1:100, m): same layout as Para%Filter
      real(c_double) :: reserved_d(8)
   end type ocfd_config

   interface
      integer(c_int) function ocfd_create(cfg, device, h) bind(C, name="ocfd_create")
         import :: c_int, c_ptr, ocfd_config
         type(ocfd_config), intent(in) :: cfg
         integer(c_int), value :: device
         type(c_ptr), intent(out) :: h
      end function
      integer(c_int) function ocfd_destroy(h) bind(C, name="ocfd_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      type(c_ptr) function ocfd_last_error() bind(C, name="ocfd_last_error")
         import :: c_ptr
      end function
      integer(c_int) function ocfd_synchronize(h) bind(C, name="ocfd_synchronize")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function ocfd_upload_metrics(h, mets) bind(C, name="ocfd_upload_metrics")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         type(c_ptr), intent(in) :: mets(19)
      end function
      integer(c_int) function ocfd_upload_primitives(h, d, u, v, w, t, istep, tt) &
         bind(C, name="ocfd_upload_primitives")
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: h
         real(c_double), intent(in) :: d(*), u(*), v(*), w(*), t(*)
         integer(c_int32_t), value :: istep
         real(c_double), value :: tt
      end function
      integer(c_int) function ocfd_upload_f(h, f) bind(C, name="ocfd_upload_f")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: f(*)
      end function
      integer(c_int) function ocfd_download_f(h, f) bind(C, name="ocfd_download_f")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: f(*)
      end function
      integer(c_int) function ocfd_download_primitives(h, d, u, v, w, t) bind(C, name="ocfd_download_primitives")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: d(*), u(*), v(*), w(*), t(*)
      end function
      integer(c_int) function ocfd_save_flow_data(h, path, with_f) bind(C, name="ocfd_save_flow_data")
         import :: c_int, c_ptr, c_char, c_int32_t
         type(c_ptr), value :: h
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int32_t), value :: with_f
      end function
      integer(c_int) function ocfd_read_flow_data(h, path, has_f) bind(C, name="ocfd_read_flow_data")
         import :: c_int, c_ptr, c_char, c_int32_t
         type(c_ptr), value :: h
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int32_t), intent(out) :: has_f
      end function
      integer(c_int) function ocfd_set_bc_boundarylayer(h, aoa, inlet, upper, nonref, pertb, wall) &
         bind(C, name="ocfd_set_bc_boundarylayer")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), value :: aoa
         type(c_ptr), value :: inlet, upper, nonref, pertb, wall
      end function
      integer(c_int) function ocfd_set_bc_time_factor(h, ht) bind(C, name="ocfd_set_bc_time_factor")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), value :: ht
      end function
      integer(c_int) function ocfd_get_time(h, istep, tt) bind(C, name="ocfd_get_time")
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), intent(out) :: istep
         real(c_double), intent(out) :: tt
      end function
      integer(c_int) function ocfd_exchange_primitives(h) bind(C, name="ocfd_exchange_primitives")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function ocfd_comput_rhybrid(h, counts) bind(C, name="ocfd_comput_rhybrid")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         type(c_ptr), value :: counts
      end function
      integer(c_int) function ocfd_du_inviscous(h) bind(C, name="ocfd_du_inviscous")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function ocfd_du_inviscous_character(h) bind(C, name="ocfd_du_inviscous_character")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function ocfd_du_viscous(h) bind(C, name="ocfd_du_viscous")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function ocfd_time_adv_rk3(h, krk) bind(C, name="ocfd_time_adv_rk3")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: krk
      end function
      integer(c_int) function ocfd_save_fn(h) bind(C, name="ocfd_save_fn")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function ocfd_comput_duvwt(h) bind(C, name="ocfd_comput_duvwt")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function ocfd_bc(h) bind(C, name="ocfd_bc")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function ocfd_set_stage_fusion(h, on) bind(C, name="ocfd_set_stage_fusion")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: on
      end function
      integer(c_int) function ocfd_filtering(h) bind(C, name="ocfd_filtering")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function ocfd_advance_clock(h) bind(C, name="ocfd_advance_clock")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function ocfd_step(h, nsteps) bind(C, name="ocfd_step")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: nsteps
      end function
      integer(c_int) function ocfd_show_flow_msg(h, out) bind(C, name="ocfd_show_flow_msg")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: out(3)
      end function
      integer(c_int) function ocfd_ana_box(h, out) bind(C, name="ocfd_ana_box")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: out(20)
      end function
   end interface

contains

   ! Stop with the library's message when a call failed (the reference's
   ! error convention: print + stop).
   subroutine ocfd_check(rc, what)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: what
      character(kind=c_char), pointer :: msg(:)
      character(len=512) :: text
      integer :: i
      if (rc >= 0) return
      call c_f_pointer(ocfd_last_error(), msg, [512])
      text = ' '
      do i = 1, 512
         if (msg(i) == c_null_char) exit
         text(i:i) = msg(i)
      end do
      print *, 'libocfd_hip: ', what, ' failed: ', trim(text)
      stop
   end subroutine ocfd_check

   ! Address of the first element of an array (for ocfd_upload_metrics).
   function ocfd_addr(x) result(p)
      real(c_double), target, intent(in) :: x(*)
      type(c_ptr) :: p
      p = c_loc(x(1))
   end function ocfd_addr

end module ocfd_hip
