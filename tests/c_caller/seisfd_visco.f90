! The bind(C) mirror of sfd_visco_params and the two entry points a Fortran
! driver of forward/Q_forward binds (as in the integration notes), plus a
! program printing the type's size and field offsets in the format of
! layout_probe.c, so a test can compare the two compilers' layouts.
module seisfd
  use iso_c_binding
  implicit none
  type, bind(C) :: sfd_visco_params
     integer(c_int) :: nx, nz, nt, delay
     real(c_float)  :: dx, dz, dt, f0
     integer(c_int) :: radius
     type(c_ptr)    :: coef
     integer(c_int) :: nmech
     real(c_float)  :: q_scale, f_low, f_high
     integer(c_int) :: sponge_xlo, sponge_xhi, sponge_zlo, sponge_zhi
     real(c_float)  :: sponge_r
     integer(c_int) :: sx, sz
     real(c_float)  :: src_sigma
     integer(c_int) :: nrec
     type(c_ptr)    :: rx, rz
     integer(c_int) :: nsnap
     type(c_ptr)    :: snap_steps
     integer(c_int) :: snap_fields
     type(c_ptr)    :: qs
  end type
  interface
     integer(c_int) function sfd_context_create(device, ctx) bind(C, name="sfd_context_create")
       import :: c_int, c_ptr
       integer(c_int), value :: device
       type(c_ptr) :: ctx
     end function
     integer(c_int) function sfd_visco_forward(ctx, p, vp, vs, rho, qp, wavelet, record, snaps) &
          bind(C, name="sfd_visco_forward")
       import :: c_int, c_ptr, sfd_visco_params
       type(c_ptr), value :: ctx
       type(sfd_visco_params) :: p
       type(c_ptr), value :: vp, vs, rho, qp, wavelet, record, snaps
     end function
  end interface
end module

program layout
  use seisfd
  implicit none
  type(sfd_visco_params), target :: p
  integer(c_intptr_t) :: b
  b = transfer(c_loc(p), b)
  print '(a,i0)', 'size ', c_sizeof(p)
  call show('nx', c_loc(p%nx)); call show('nz', c_loc(p%nz)); call show('nt', c_loc(p%nt))
  call show('delay', c_loc(p%delay)); call show('dx', c_loc(p%dx)); call show('dz', c_loc(p%dz))
  call show('dt', c_loc(p%dt)); call show('f0', c_loc(p%f0)); call show('radius', c_loc(p%radius))
  call show('coef', c_loc(p%coef)); call show('nmech', c_loc(p%nmech)); call show('q_scale', c_loc(p%q_scale))
  call show('f_low', c_loc(p%f_low)); call show('f_high', c_loc(p%f_high))
  call show('sponge_xlo', c_loc(p%sponge_xlo)); call show('sponge_xhi', c_loc(p%sponge_xhi))
  call show('sponge_zlo', c_loc(p%sponge_zlo)); call show('sponge_zhi', c_loc(p%sponge_zhi))
  call show('sponge_r', c_loc(p%sponge_r)); call show('sx', c_loc(p%sx)); call show('sz', c_loc(p%sz))
  call show('src_sigma', c_loc(p%src_sigma)); call show('nrec', c_loc(p%nrec)); call show('rx', c_loc(p%rx))
  call show('rz', c_loc(p%rz)); call show('nsnap', c_loc(p%nsnap)); call show('snap_steps', c_loc(p%snap_steps))
  call show('snap_fields', c_loc(p%snap_fields)); call show('qs', c_loc(p%qs))
contains
  subroutine show(name, q)
    character(*), intent(in) :: name
    type(c_ptr), intent(in) :: q
    integer(c_intptr_t) :: a
    a = transfer(q, a)
    print '(a,a,i0)', name, ' ', a - b
  end subroutine
end program
