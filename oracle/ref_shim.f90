! ref_shim.f90 -- bind(C) entry points over the reference's own Fortran modules, used
! only to pin the oracle (test infrastructure).  Built by oracle/Makefile together with
! the reference sources tensor_element.f90, random.f90, io.f90 and transform.f90 taken
! from the reference tree in place; nothing of them is copied here.
module mrm_ref_shim
  use iso_c_binding
  use mselem_m, only: volume_averaged_n, plane_to_plane
  use random_m, only: random_s, random_create, random_uniform, random_normal3d
  implicit none
  type gen_holder
    type(random_s), pointer :: g => null()
  end type
  type(gen_holder), save :: gens(64)
contains
  ! volume_averaged_n (tensor_element.f90:13-102): n(3,3) accumulates -t
  function ref_volume_averaged_n(rx,ry,rz,xs,ys,zs,xo,yo,zo,n) result(vol) bind(C,name='ref_volume_averaged_n')
    real(c_double), value :: rx,ry,rz,xs,ys,zs,xo,yo,zo
    real(c_double), intent(inout) :: n(3,3)
    real(c_double) :: vol
    vol = volume_averaged_n(rx,ry,rz,xs,ys,zs,xo,yo,zo,n)
  end function
  ! plane_to_plane (tensor_element.f90:219-261): out(nx,ny,3,3)
  subroutine ref_plane_to_plane(nx,ny,dx,dy,dz,dzo,rz,imx,imy,out) bind(C,name='ref_plane_to_plane')
    integer(c_int), value :: nx,ny,imx,imy
    real(c_double), value :: dx,dy,dz,dzo,rz
    real(c_double), intent(out) :: out(nx,ny,3,3)
    out = plane_to_plane(nx,ny,dx,dy,dz,dzo,rz,imx,imy)
  end subroutine
  ! random_create (random.f90:166-183) into slot h (1..64)
  subroutine ref_random_create(h, seed) bind(C,name='ref_random_create')
    integer(c_int), value :: h, seed
    gens(h)%g => random_create('shim', seed)
  end subroutine
  subroutine ref_random_uniform(h, n, out) bind(C,name='ref_random_uniform')
    integer(c_int), value :: h, n
    real(c_double), intent(out) :: out(n)
    integer :: i
    do i = 1, n
      out(i) = random_uniform(gens(h)%g)
    end do
  end subroutine
  subroutine ref_random_normal3d(h, n, out) bind(C,name='ref_random_normal3d')
    integer(c_int), value :: h, n
    real(c_double), intent(out) :: out(3,n)
    integer :: i
    do i = 1, n
      out(:,i) = random_normal3d(gens(h)%g)
    end do
  end subroutine
end module
