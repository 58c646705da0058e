! Probe driver for the compiled reference solver (test infrastructure only).
!
! Links against the reference's own object files (built from the sources under
! the reference tree by oracle/Makefile) and calls its routines in the order of
! one NS_solver time step (OCFD_NS_Solver.f90:21-78), dumping every intermediate
! array as raw little-endian float64 (Fortran column-major order, halos included
! where the reference keeps them) into the directory given as argument 1.
! Argument 2 (optional) is the number of whole steps to run after the dumped
! first step; the state after those steps is dumped too.
!
! Files written (NAME.bin):
!   metrics: Akx..Asz, Ajac, Akx1..Asz1 (halo'd, after init_mesh)
!   s0_{d,u,v,w,T}     primitives after init + OCFD_bc, after the stage-1 halo exchange
!   s0_f               conserved variables at step start
!   rhyb               Rhybrid after comput_Rhybrid (only when HYBRID)
!   k1_duinv           du after du_inviscous / du_inviscous_Character (stage 1)
!   k1_dutot           du after du_viscous (stage 1)
!   k1_f               f after OCFD_time_adv_RK3(1)
!   k1_{d,u,v,w,T}     primitives after comput_duvwT + OCFD_bc (stage 1)
!   k3_f_prefilter     f after stage 3, before filtering
!   n1_f, n1_{d..T}    state after the first whole step (filter included)
!   nN_f, nN_{d..T}    state after 1+N whole steps
!   hist.txt           per step: Istep tt E K S (show_flow_msg sums) KE enstrophy
program ocfd_probe
   use flow_data
   use viscous_data
   implicit none
   integer:: ierr, KRK, nextra, n, nargs
   character(len=512):: outdir, arg
   integer, parameter:: IBUFFER_SIZE = 1000000
   real(kind=8):: BUFFER_MPI(IBUFFER_SIZE)

   call mpi_init(ierr)
   call mpi_comm_rank(MPI_COMM_WORLD, my_id, ierr)
   call MPI_BUFFER_ATTACH(BUFFER_MPI, 8*IBUFFER_SIZE, ierr)
   nargs = command_argument_count()
   outdir = "."
   nextra = 0
   if (nargs >= 1) call get_command_argument(1, outdir)
   if (nargs >= 2) then
      call get_command_argument(2, arg)
      read (arg, *) nextra
   end if

   call read_parameter
   call partation3d_mpi
   call set_parameters
   if (np_size /= 1) then
      print *, "ocfd_probe runs on one rank only"
      stop
   end if

   call allocate_flow_data
   call allocate_inviscous_data
   call allocate_vicous_data
   call init3d
   call OCFD_bc

   call dump3h('Akx', Akx); call dump3h('Aky', Aky); call dump3h('Akz', Akz)
   call dump3h('Aix', Aix); call dump3h('Aiy', Aiy); call dump3h('Aiz', Aiz)
   call dump3h('Asx', Asx); call dump3h('Asy', Asy); call dump3h('Asz', Asz)
   call dump3h('Ajac', Ajac)
   call dump3h('Akx1', Akx1); call dump3h('Aky1', Aky1); call dump3h('Akz1', Akz1)
   call dump3h('Aix1', Aix1); call dump3h('Aiy1', Aiy1); call dump3h('Aiz1', Aiz1)
   call dump3h('Asx1', Asx1); call dump3h('Asy1', Asy1); call dump3h('Asz1', Asz1)

   open (91, file=trim(outdir)//'/hist.txt', status='replace')

   ! ---- first step, instrumented (same call order as the driver loop) ----
   fn = f
   call dump4('s0_f', f)
   do KRK = 1, 3
      call exchange_boundary_xyz(d)
      call exchange_boundary_xyz(u)
      call exchange_boundary_xyz(v)
      call exchange_boundary_xyz(w)
      call exchange_boundary_xyz(T)
      if (KRK == 1) then
         call dump3h('s0_d', d); call dump3h('s0_u', u); call dump3h('s0_v', v)
         call dump3h('s0_w', w); call dump3h('s0_T', T)
      end if
      if (KRK .eq. 1 .and. Scheme%Scheme_Invis == OCFD_Scheme_Hybrid) then
         call comput_Rhybrid
         call dump3h('rhyb', Rhybrid)
      end if
      if (Para%IF_Scheme_Character .eq. 0) then
         call du_inviscous
      else
         call du_inviscous_Character
      end if
      if (KRK == 1) call dump4('k1_duinv', du)
      if (Para%IF_Viscous .eq. 1) call du_viscous
      if (KRK == 1) call dump4('k1_dutot', du)
      call OCFD_time_adv_RK3(KRK)
      if (KRK == 1) call dump4('k1_f', f)
      call comput_duvwT
      call OCFD_bc
      if (KRK == 1) then
         call dump3h('k1_d', d); call dump3h('k1_u', u); call dump3h('k1_v', v)
         call dump3h('k1_w', w); call dump3h('k1_T', T)
      end if
   end do
   call dump4('k3_f_prefilter', f)
   Istep = Istep + 1
   tt = tt + Para%dt
   call filtering(f)
   call dump4('n1_f', f)
   call dump3h('n1_d', d); call dump3h('n1_u', u); call dump3h('n1_v', v)
   call dump3h('n1_w', w); call dump3h('n1_T', T)
   call write_hist

   ! ---- further whole steps, not instrumented ----
   do n = 1, nextra
      fn = f
      do KRK = 1, 3
         call exchange_boundary_xyz(d)
         call exchange_boundary_xyz(u)
         call exchange_boundary_xyz(v)
         call exchange_boundary_xyz(w)
         call exchange_boundary_xyz(T)
         if (KRK .eq. 1 .and. Scheme%Scheme_Invis == OCFD_Scheme_Hybrid) call comput_Rhybrid
         if (Para%IF_Scheme_Character .eq. 0) then
            call du_inviscous
         else
            call du_inviscous_Character
         end if
         if (Para%IF_Viscous .eq. 1) call du_viscous
         call OCFD_time_adv_RK3(KRK)
         call comput_duvwT
         call OCFD_bc
      end do
      Istep = Istep + 1
      tt = tt + Para%dt
      call filtering(f)
      call write_hist
   end do
   if (nextra > 0) then
      call dump4('nN_f', f)
      call dump3h('nN_d', d); call dump3h('nN_u', u); call dump3h('nN_v', v)
      call dump3h('nN_w', w); call dump3h('nN_T', T)
   end if
   close (91)
   call mpi_finalize(ierr)

contains

   subroutine dump3h(name, a)
      character(len=*), intent(in):: name
      real(kind=OCFD_REAL_KIND), intent(in):: a(1 - LAP:nx + LAP, 1 - LAP:ny + LAP, 1 - LAP:nz + LAP)
      open (77, file=trim(outdir)//'/'//name//'.bin', form='unformatted', access='stream', status='replace')
      write (77) a
      close (77)
   end subroutine

   subroutine dump4(name, a)
      character(len=*), intent(in):: name
      real(kind=OCFD_REAL_KIND), intent(in):: a(nx, ny, nz, 5)
      open (77, file=trim(outdir)//'/'//name//'.bin', form='unformatted', access='stream', status='replace')
      write (77) a
      close (77)
   end subroutine

   ! Volume-weighted sums of show_flow_msg (OCFD_Time_Adv.f90:77-89) and the
   ! kinetic energy / enstrophy of ana_box (OCFD_ANA.f90:200-262), both divided
   ! by the global cell count as the reference does.  The velocity gradients go
   ! through the reference's own OCFD_dx0/dy0/dz0 after a halo exchange.
   subroutine write_hist
      real(kind=8):: E(3), ke, ens, p00, tmp0, tmp1, ux, uy, uz, vx, vy, vz, wx, wy, wz
      integer:: i, j, k
      p00 = 1.d0/(Para%gamma*Para%Ma*Para%Ma)
      tmp1 = 1.d0 - Para%gamma
      E = 0.d0
      ke = 0.d0
      ens = 0.d0
      do k = 1, nz
      do j = 1, ny
      do i = 1, nx
         tmp0 = 1.d0/Ajac(i, j, k)
         E(1) = E(1) + f(i, j, k, 5)*tmp0
         E(2) = E(2) + d(i, j, k)*(u(i, j, k)*u(i, j, k) + v(i, j, k)*v(i, j, k) + w(i, j, k)*w(i, j, k))*0.5d0*tmp0
         E(3) = E(3) + p00*d(i, j, k)**tmp1*T(i, j, k)*tmp0
      end do
      end do
      end do
      call exchange_boundary_xyz(u)
      call exchange_boundary_xyz(v)
      call exchange_boundary_xyz(w)
      call OCFD_dx0(u, uk, Scheme%Scheme_Vis); call OCFD_dy0(u, ui, Scheme%Scheme_Vis); call OCFD_dz0(u, us, Scheme%Scheme_Vis)
      call OCFD_dx0(v, vk, Scheme%Scheme_Vis); call OCFD_dy0(v, vi, Scheme%Scheme_Vis); call OCFD_dz0(v, vs, Scheme%Scheme_Vis)
      call OCFD_dx0(w, wk, Scheme%Scheme_Vis); call OCFD_dy0(w, wi, Scheme%Scheme_Vis); call OCFD_dz0(w, ws, Scheme%Scheme_Vis)
      do k = 1, nz
      do j = 1, ny
      do i = 1, nx
         ux = uk(i, j, k)*Akx(i, j, k) + ui(i, j, k)*Aix(i, j, k) + us(i, j, k)*Asx(i, j, k)
         uy = uk(i, j, k)*Aky(i, j, k) + ui(i, j, k)*Aiy(i, j, k) + us(i, j, k)*Asy(i, j, k)
         uz = uk(i, j, k)*Akz(i, j, k) + ui(i, j, k)*Aiz(i, j, k) + us(i, j, k)*Asz(i, j, k)
         vx = vk(i, j, k)*Akx(i, j, k) + vi(i, j, k)*Aix(i, j, k) + vs(i, j, k)*Asx(i, j, k)
         vy = vk(i, j, k)*Aky(i, j, k) + vi(i, j, k)*Aiy(i, j, k) + vs(i, j, k)*Asy(i, j, k)
         vz = vk(i, j, k)*Akz(i, j, k) + vi(i, j, k)*Aiz(i, j, k) + vs(i, j, k)*Asz(i, j, k)
         wx = wk(i, j, k)*Akx(i, j, k) + wi(i, j, k)*Aix(i, j, k) + ws(i, j, k)*Asx(i, j, k)
         wy = wk(i, j, k)*Aky(i, j, k) + wi(i, j, k)*Aiy(i, j, k) + ws(i, j, k)*Asy(i, j, k)
         wz = wk(i, j, k)*Akz(i, j, k) + wi(i, j, k)*Aiz(i, j, k) + ws(i, j, k)*Asz(i, j, k)
         ke = ke + (u(i, j, k)*u(i, j, k) + v(i, j, k)*v(i, j, k) + w(i, j, k)*w(i, j, k))*d(i, j, k)*0.5d0
         ens = ens + (wy - vz)**2 + (uz - wx)**2 + (vx - uy)**2
      end do
      end do
      end do
      E = E/(nx_global*ny_global*nz_global)
      ke = ke/(nx_global*ny_global*nz_global)
      ens = 0.5d0*ens/(nx_global*ny_global*nz_global)
      write (91, '(I8,1X,6(ES25.17E3,1X))') Istep, tt, E(1), E(2), E(3), ke, ens
   end subroutine

end program ocfd_probe
