! ref_shim.F90 -- C-callable entry points over the reference's own Fortran
! module `Tomography` (tomography/tv_hpc.F90:1-390), compiled unmodified
! from the reference tree by oracle/Makefile (target `ref`) into
! oracle/_ref/libacumen_ref.so.
!
! TEST INFRASTRUCTURE ONLY.  The test-suite uses it to check the C oracle
! against the reference's real FP32 arithmetic; nothing in the product
! links or loads it.
!
! The module keeps its state in public globals (tv_hpc.F90:3-7) and is
! driven the way the reference's drivers drive it (tv_hpc.F90:546-562,
! 609-620): set nimage, ntheta, wpu; call int_radon; fill theta; fill f;
! call tvopt / get_slice.  Its int_radon allocates every module array and
! fin_radon frees nothing (tv_hpc.F90:80-82), so ref_init releases the
! arrays of a previous set-up before calling int_radon again.
! Arrays cross the boundary in the reference's layout: image (x, y) and
! sinogram (detector, angle), first index fastest.
module acumen_ref_shim
  use iso_c_binding
  use Tomography
  implicit none
contains

  subroutine release()
    if (allocated(filter)) deallocate(filter)
    if (allocated(theta)) deallocate(theta)
    if (allocated(cmask)) deallocate(cmask)
    if (allocated(img_rec)) deallocate(img_rec)
    if (allocated(sinogramCS)) deallocate(sinogramCS)
    if (allocated(uk)) deallocate(uk)
    if (allocated(ukp)) deallocate(ukp)
    if (allocated(dk_x)) deallocate(dk_x)
    if (allocated(dk_y)) deallocate(dk_y)
    if (allocated(bk_x)) deallocate(bk_x)
    if (allocated(bk_y)) deallocate(bk_y)
    if (allocated(ifkt)) deallocate(ifkt)
    if (allocated(rhs)) deallocate(rhs)
    if (allocated(iukt)) deallocate(iukt)
    if (allocated(r)) deallocate(r)
    if (allocated(p)) deallocate(p)
    if (allocated(irpt)) deallocate(irpt)
    if (allocated(Ap)) deallocate(Ap)
    if (allocated(f)) deallocate(f)
    if (allocated(fk)) deallocate(fk)
    if (allocated(ruk)) deallocate(ruk)
    if (allocated(rp)) deallocate(rp)
  end subroutine release

  ! int_radon (tv_hpc.F90:11-78) for an N x N image and T angles, then the
  ! angle fill of the drivers (tv_hpc.F90:560-562).
  subroutine ref_init(n, t, w, th) bind(C, name="ref_init")
    integer(c_int), value :: n, t, w
    real(c_float), intent(in) :: th(t)
    call release()
    nimage = n
    ntheta = t
    wpu = w
    call int_radon()
    theta(:) = th(:)
  end subroutine ref_init

  ! filter (2N+1 taps) and cmask (N x N) as int_radon left them.
  subroutine ref_tables(h, cm) bind(C, name="ref_tables")
    real(c_float), intent(out) :: h(nfilt), cm(nimage, nimage)
    h(:) = filter(:)
    cm(:, :) = cmask(:, :)
  end subroutine ref_tables

  ! radon(sinogram, img) (tv_hpc.F90:129-159) through the module arrays the
  ! reference itself passes (uk in, ruk out).
  subroutine ref_radon(img, sino) bind(C, name="ref_radon")
    real(c_float), intent(in) :: img(nimage, nimage)
    real(c_float), intent(out) :: sino(nimage, ntheta)
    uk(:, :) = img(:, :)
    call radon(ruk, uk)
    sino(:, :) = ruk(:, :)
  end subroutine ref_radon

  ! iradon(img_rec, sinogram) (tv_hpc.F90:84-126) through fk in, ifkt out.
  subroutine ref_iradon(sino, img) bind(C, name="ref_iradon")
    real(c_float), intent(in) :: sino(nimage, ntheta)
    real(c_float), intent(out) :: img(nimage, nimage)
    fk(:, :) = sino(:, :)
    call iradon(ifkt, fk)
    img(:, :) = ifkt(:, :)
  end subroutine ref_iradon

  ! The filtered sinogram iradon leaves in sinogramCS (tv_hpc.F90:93-105).
  subroutine ref_filtered(out) bind(C, name="ref_filtered")
    real(c_float), intent(out) :: out(nimage, ntheta)
    out(:, :) = sinogramCS(:, :)
  end subroutine ref_filtered

  ! tvopt (tv_hpc.F90:162-363): f is the windowed sinogram, the module
  ! parameters are set as the drivers set them (tv_hpc.F90:483-486).
  subroutine ref_tvopt(sino, mu_, lambda_, nouter_, ninner_, ncg_, img_tv, img_fbp) &
      bind(C, name="ref_tvopt")
    real(c_float), intent(in) :: sino(nimage, ntheta)
    real(c_float), value :: mu_, lambda_
    integer(c_int), value :: nouter_, ninner_, ncg_
    real(c_float), intent(out) :: img_tv(nimage, nimage), img_fbp(nimage, nimage)
    f(:, :) = sino(:, :)
    mu = mu_
    lambda = lambda_
    nouter = nouter_
    ninnner = ninner_
    ncg = ncg_
    call tvopt(img_tv, img_fbp, 1)
  end subroutine ref_tvopt

  ! get_slice (tv_hpc.F90:365-387): FBP of f.
  subroutine ref_get_slice(sino, img_fbp) bind(C, name="ref_get_slice")
    real(c_float), intent(in) :: sino(nimage, ntheta)
    real(c_float), intent(out) :: img_fbp(nimage, nimage)
    f(:, :) = sino(:, :)
    call get_slice(img_fbp)
  end subroutine ref_get_slice

  subroutine ref_fin() bind(C, name="ref_fin")
    call fin_radon()
    call release()
  end subroutine ref_fin

end module acumen_ref_shim
