! User hooks linked into the probe driver in place of the reference's empty
! OCFD_User_Def.f90 (test infrastructure only).  ocfd_bc_User is the reference's
! designated extension point for a case-specific boundary condition
! (OCFD_BC.f90:18-19).  This one implements the isothermal-wall / far-field
! pair used by the curvilinear O-grid case:
!   j=1  : no-slip wall, T = BC_para(1), density from a second-order
!          extrapolation of the pressure p ~ d*T along j
!   j=ny : free stream d=1, u=1, v=w=0, T=1
! BC_para(1) <= 0 turns the wall into an adiabatic one (T extrapolated).
subroutine ocfd_bc_User
   use flow_data
   implicit none
   integer:: i, k
   real(kind=OCFD_REAL_KIND):: Tw
   Tw = Para%BC_para(1)
   if (npy .eq. 0) then
      do k = 1, nz
      do i = 1, nx
         u(i, 1, k) = 0.d0
         v(i, 1, k) = 0.d0
         w(i, 1, k) = 0.d0
         if (Tw > 0.d0) then
            d(i, 1, k) = (4.d0*d(i, 2, k)*T(i, 2, k) - d(i, 3, k)*T(i, 3, k))/(3.d0*Tw)
            T(i, 1, k) = Tw
         else
            d(i, 1, k) = (4.d0*d(i, 2, k) - d(i, 3, k))/3.d0
            T(i, 1, k) = (4.d0*T(i, 2, k) - T(i, 3, k))/3.d0
         end if
      end do
      end do
   end if
   if (npy .eq. npy0 - 1) then
      do k = 1, nz
      do i = 1, nx
         d(i, ny, k) = 1.d0
         u(i, ny, k) = 1.d0
         v(i, ny, k) = 0.d0
         w(i, ny, k) = 0.d0
         T(i, ny, k) = 1.d0
      end do
      end do
   end if
end

subroutine ana_user(ana_data)
   use flow_data
   implicit none
   real(kind=OCFD_REAL_KIND):: ana_data(50)
end

subroutine OCFD_Scheme_USER_P(v, hh, nx, LAP, ib, ie)
   use OCFD_constants
   implicit none
   integer nx, LAP, ib, ie
   real(kind=OCFD_REAL_KIND):: v(1 - LAP:nx + LAP), hh(0:nx)
   stop 'user scheme not provided'
end

subroutine OCFD_Scheme_USER_M(v, hh, nx, LAP, ib, ie)
   use OCFD_constants
   implicit none
   integer nx, LAP, ib, ie
   real(kind=OCFD_REAL_KIND):: v(1 - LAP:nx + LAP), hh(0:nx)
   stop 'user scheme not provided'
end

subroutine Stencil_Scheme_User(Ka_P, Kb_P, Ka_M, Kb_M)
   implicit none
   integer Ka_P, Kb_P, Ka_M, Kb_M
   stop 'user scheme not provided'
end

subroutine hh_Scheme_USER_P(Ka, Kb, v, hh)
   use OCFD_constants
   implicit none
   integer Ka, Kb
   real(kind=OCFD_REAL_KIND):: v(Ka:Kb), hh
   stop 'user scheme not provided'
end

subroutine hh_Scheme_USER_M(Ka, Kb, v, hh)
   use OCFD_constants
   implicit none
   integer Ka, Kb
   real(kind=OCFD_REAL_KIND):: v(Ka:Kb), hh
   stop 'user scheme not provided'
end
