! Test-infrastructure driver: C-callable entry points into the reference
! BRAIDS-HKT Fortran solver (cutcell_solver.f90 and the vendored GeoClaw
! Riemann routines), compiled from the sources where they lie.  Used only to
! pin the oracle (oracle/swe_oracle.c); never part of the product.
module braids_ref_driver
 use iso_c_binding
 use geoclaw_module, only: grav,dry_tolerance
 use cutcell_solver, only: wave_flux
 use braids_bedload_mpm_module, only: braids_mpm_bedload
 implicit none
contains
 subroutine ref_set_constants(g,dry) bind(C,name='ref_set_constants')
  real(c_double),value :: g,dry
  grav=g;dry_tolerance=dry
 end subroutine

 ! wave_flux, cutcell_solver.f90:9135-9159
 subroutine ref_wave_flux(left,right,bl,br,n,fl,fr,speed,rightmost) bind(C,name='ref_wave_flux')
  real(c_double),intent(in) :: left(3),right(3),n(2)
  real(c_double),value :: bl,br
  real(c_double),intent(out) :: fl(3),fr(3),speed,rightmost
  call wave_flux(left,right,bl,br,n,fl,fr,speed,rightmost)
 end subroutine

 ! riemanntype, geoclaw_riemann_utils.f:547-658
 subroutine ref_riemanntype(hl,hr,ul,ur,hm,s1m,s2m,rare1,rare2) bind(C,name='ref_riemanntype')
  real(c_double),value :: hl,hr,ul,ur
  real(c_double),intent(out) :: hm,s1m,s2m
  integer(c_int),intent(out) :: rare1,rare2
  logical :: r1,r2
  external riemanntype
  call riemanntype(hl,hr,ul,ur,hm,s1m,s2m,r1,r2,1,dry_tolerance,grav)
  rare1=merge(1,0,r1);rare2=merge(1,0,r2)
 end subroutine

 ! Meyer-Peter--Mueller bedload per unit width, braids_bedload_mpm_module.f90:15-41
 subroutine ref_mpm_bedload(u,v,cf,d50,rd,g,thc,qx,qy,valid) bind(C,name='ref_mpm_bedload')
  real(c_double),value :: u,v,cf,d50,rd,g,thc
  real(c_double),intent(out) :: qx,qy
  integer(c_int),intent(out) :: valid
  logical :: ok
  call braids_mpm_bedload(u,v,cf,d50,rd,g,thc,qx,qy,ok)
  valid=merge(1,0,ok)
 end subroutine
end module
