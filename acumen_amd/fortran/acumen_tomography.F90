! acumen_tomography.F90 -- drop-in replacement of Acumen's Fortran module
! `Tomography` (tomography/tv_hpc.F90:1-390; copies in fbp_hpc.F90:1-154 and
! tv_hpc_slice.F90:1-366) on top of the C ABI of libacumen_hip.so
! (include/acumen.h).
!
! Same public globals and subroutine names, argument lists and calling order
! as the reference: the driver sets nimage, ntheta, wpu (and mu, lambda,
! nouter, ninnner, ncg for TV), calls int_radon(), fills theta(:) and f(:,:),
! then calls get_slice / tvopt / radon / iradon, and finally fin_radon().
! Arrays are FP32, first index fastest, exactly the reference's layout.
!
! Differences: the state lives on the GPU behind an opaque handle; every
! library status is checked and a failure stops the program with the
! library's message (the reference checks no status at all).
! batch_* procedures extend the module with many slices per launch.
module Tomography
  use iso_c_binding
  implicit none

  integer, public :: ncg, ninnner, nouter, nimage, ntheta, nfilt, nstart, wpu, add_Nx
  real, public :: pi, midpoint, lambda, mu
  real, public, dimension(:), allocatable, target :: theta
  real, public, dimension(:,:), allocatable, target :: f

  type(c_ptr), private, save :: ctx = c_null_ptr, plan = c_null_ptr
  type(c_ptr), private, save :: bat = c_null_ptr
  integer, private, save :: bat_n = 0

  integer(c_int), parameter, public :: ACU_IMG_TV = 0, ACU_IMG_FBP = 1
  integer(c_int), parameter, public :: ACU_TV_PLAIN = 0, ACU_TV_BATCHED = 1, ACU_TV_GRAPH = 2

  interface
    function acu_last_error() bind(C, name="acu_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function acu_ctx_create(device, out) bind(C, name="acu_ctx_create") result(rc)
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: out
      integer(c_int) :: rc
    end function
    function acu_ctx_destroy(c) bind(C, name="acu_ctx_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int) :: rc
    end function
    function acu_plan_create(c, n, t, w, th, out) bind(C, name="acu_plan_create") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: c, th
      integer(c_int), value :: n, t, w
      type(c_ptr) :: out
      integer(c_int) :: rc
    end function
    function acu_batch_create(c, ns, n, t, w, th, out) bind(C, name="acu_batch_create") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: ns
      integer(c_int) :: n(*), t(*), w(*)
      type(c_ptr) :: th(*)
      type(c_ptr) :: out
      integer(c_int) :: rc
    end function
    function acu_batch_destroy(b) bind(C, name="acu_batch_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: rc
    end function
    function acu_batch_set_theta(b, s, th) bind(C, name="acu_batch_set_theta") result(rc)
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: b
      integer(c_int), value :: s
      real(c_float) :: th(*)
      integer(c_int) :: rc
    end function
    function acu_batch_set_sino(b, s, fh) bind(C, name="acu_batch_set_sino") result(rc)
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: b
      integer(c_int), value :: s
      real(c_float) :: fh(*)
      integer(c_int) :: rc
    end function
    function acu_batch_get_image(b, s, which, img) bind(C, name="acu_batch_get_image") result(rc)
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: b
      integer(c_int), value :: s, which
      real(c_float) :: img(*)
      integer(c_int) :: rc
    end function
    function acu_batch_get_delta(b, s, d, n) bind(C, name="acu_batch_get_delta") result(rc)
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: b
      integer(c_int), value :: s, n
      real(c_float) :: d(*)
      integer(c_int) :: rc
    end function
    function acu_batch_fbp(b) bind(C, name="acu_batch_fbp") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: rc
    end function
    function acu_batch_tvopt(b, mu_, lam, no, ni, nc, mode) bind(C, name="acu_batch_tvopt") result(rc)
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: b
      real(c_float), value :: mu_, lam
      integer(c_int), value :: no, ni, nc, mode
      integer(c_int) :: rc
    end function
    function acu_radon(b, sino, img) bind(C, name="acu_radon") result(rc)
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: b
      real(c_float) :: sino(*), img(*)
      integer(c_int) :: rc
    end function
    function acu_iradon(b, img, sino) bind(C, name="acu_iradon") result(rc)
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: b
      real(c_float) :: img(*), sino(*)
      integer(c_int) :: rc
    end function
    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  subroutine acu_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    n = int(c_strlen(acu_last_error()))
    call c_f_pointer(acu_last_error(), msg, [n])
    write(*,*) "acumen: ", what, " failed (", rc, "): ", msg(1:n)
    error stop 2
  end subroutine acu_check

  subroutine ensure_ctx()
    if (.not. c_associated(ctx)) call acu_check(acu_ctx_create(0_c_int, ctx), "acu_ctx_create")
  end subroutine ensure_ctx

  ! int_radon (tv_hpc.F90:11-78): derived scalars, device state, theta and f
  subroutine int_radon()
    midpoint = (nimage + 1) / 2
    pi = 4.0 * atan(1.0)
    nfilt = 2 * nimage + 1
    nstart = nimage
    call ensure_ctx()
    if (c_associated(plan)) call acu_check(acu_batch_destroy(plan), "acu_batch_destroy")
    call acu_check(acu_plan_create(ctx, int(nimage, c_int), int(ntheta, c_int), int(wpu, c_int), &
                                   c_null_ptr, plan), "acu_plan_create")
    if (allocated(theta)) deallocate(theta)
    if (allocated(f)) deallocate(f)
    allocate(theta(ntheta))
    allocate(f(nimage, ntheta))
    theta = 0.0
    f = 0.0
  end subroutine int_radon

  ! fin_radon (tv_hpc.F90:80-82): here it releases the device state
  subroutine fin_radon()
    if (c_associated(plan)) call acu_check(acu_batch_destroy(plan), "acu_batch_destroy")
    plan = c_null_ptr
    call batch_free()
  end subroutine fin_radon

  ! iradon(img_rec, sinogram) (tv_hpc.F90:84-126)
  subroutine iradon(img_rec, sinogram)
    real, intent(in) :: sinogram(nimage, ntheta)
    real, intent(out) :: img_rec(nimage, nimage)
    call acu_check(acu_batch_set_theta(plan, 0_c_int, theta), "set theta")
    call acu_check(acu_iradon(plan, img_rec, sinogram), "iradon")
  end subroutine iradon

  ! radon(sinogram, img) (tv_hpc.F90:129-159)
  subroutine radon(sinogram, img)
    real, intent(in) :: img(nimage, nimage)
    real, intent(out) :: sinogram(nimage, ntheta)
    call acu_check(acu_batch_set_theta(plan, 0_c_int, theta), "set theta")
    call acu_check(acu_radon(plan, sinogram, img), "radon")
  end subroutine radon

  ! tvopt(img_rec_tv, img_rec_fbp, jslice) (tv_hpc.F90:162-363) on module array f
  subroutine tvopt(img_rec_tv, img_rec_fbp, jslice)
    real, intent(out) :: img_rec_tv(nimage, nimage), img_rec_fbp(nimage, nimage)
    integer, intent(in) :: jslice
    real, allocatable :: delta(:)
    integer :: jouter
    call acu_check(acu_batch_set_theta(plan, 0_c_int, theta), "set theta")
    call acu_check(acu_batch_set_sino(plan, 0_c_int, f), "set f")
    call acu_check(acu_batch_tvopt(plan, real(mu, c_float), real(lambda, c_float), int(nouter, c_int), &
                                   int(ninnner, c_int), int(ncg, c_int), ACU_TV_GRAPH), "tvopt")
    call acu_check(acu_batch_get_image(plan, 0_c_int, ACU_IMG_TV, img_rec_tv), "get tv image")
    call acu_check(acu_batch_get_image(plan, 0_c_int, ACU_IMG_FBP, img_rec_fbp), "get fbp image")
    if (nouter > 0) then
      allocate(delta(nouter))
      call acu_check(acu_batch_get_delta(plan, 0_c_int, delta, int(nouter, c_int)), "get delta")
      do jouter = 1, nouter
        if (mod(jouter, 50) .EQ. 0) write(*,*) "Slice:", jslice, "Loop: ", jouter, "delta:", delta(jouter)
      end do
      deallocate(delta)
    end if
  end subroutine tvopt

  ! get_slice(img_rec_fbp) (fbp_hpc.F90:129-150): FBP of module array f
  subroutine get_slice(img_rec_fbp)
    real, intent(out) :: img_rec_fbp(nimage, nimage)
    call acu_check(acu_batch_set_theta(plan, 0_c_int, theta), "set theta")
    call acu_check(acu_batch_set_sino(plan, 0_c_int, f), "set f")
    call acu_check(acu_batch_fbp(plan), "fbp")
    call acu_check(acu_batch_get_image(plan, 0_c_int, ACU_IMG_FBP, img_rec_fbp), "get fbp image")
  end subroutine get_slice

  ! ---- batched extension: nb slices of the module's (nimage, ntheta, wpu, theta) ----

  subroutine batch_free()
    if (c_associated(bat)) call acu_check(acu_batch_destroy(bat), "acu_batch_destroy")
    bat = c_null_ptr
    bat_n = 0
  end subroutine batch_free

  ! Device state for nb slices that share the module's geometry.
  subroutine batch_init(nb)
    integer, intent(in) :: nb
    integer(c_int), allocatable :: ns(:), ts(:), ws(:)
    type(c_ptr), allocatable :: th(:)
    integer :: i
    call ensure_ctx()
    if (c_associated(bat) .and. bat_n == nb) return
    call batch_free()
    allocate(ns(nb), ts(nb), ws(nb), th(nb))
    ns = int(nimage, c_int)
    ts = int(ntheta, c_int)
    ws = int(wpu, c_int)
    do i = 1, nb
      th(i) = c_loc(theta)
    end do
    call acu_check(acu_batch_create(ctx, int(nb, c_int), ns, ts, ws, th, bat), "acu_batch_create")
    bat_n = nb
  end subroutine batch_init

  ! f(:,:) of batch slot i (1-based)
  subroutine batch_set_slice(i, fslice)
    integer, intent(in) :: i
    real, intent(in) :: fslice(nimage, ntheta)
    call acu_check(acu_batch_set_theta(bat, int(i - 1, c_int), theta), "set theta")
    call acu_check(acu_batch_set_sino(bat, int(i - 1, c_int), fslice), "set f")
  end subroutine batch_set_slice

  subroutine batch_fbp()
    call acu_check(acu_batch_fbp(bat), "batch fbp")
  end subroutine batch_fbp

  subroutine batch_tvopt()
    call acu_check(acu_batch_tvopt(bat, real(mu, c_float), real(lambda, c_float), int(nouter, c_int), &
                                   int(ninnner, c_int), int(ncg, c_int), ACU_TV_GRAPH), "batch tvopt")
  end subroutine batch_tvopt

  subroutine batch_get(i, which, img)
    integer, intent(in) :: i
    integer(c_int), intent(in) :: which
    real, intent(out) :: img(nimage, nimage)
    call acu_check(acu_batch_get_image(bat, int(i - 1, c_int), which, img), "batch get image")
  end subroutine batch_get

  subroutine batch_delta(i, delta)
    integer, intent(in) :: i
    real, intent(out) :: delta(nouter)
    call acu_check(acu_batch_get_delta(bat, int(i - 1, c_int), delta, int(nouter, c_int)), "delta")
  end subroutine batch_delta

end module Tomography
