! projections_to_sinogram_hip.F90 -- corrected sinograms from raw projections
! on one GPU: the driver of tomography/projections_to_sinogram_hpc.F90
! without MPI.
!
! Kept from the reference: environment ROOT_CT_DIR / EXP_CT_NAME, the header
! of fileinfo_projections.dat (offsets 0-3: nfiles, nimage, n_files_sino,
! n_files_rec, :56-71), theta.dat, the projection -> file/offset map of
! projections{js}.dat (:146-148, 391-407), the 0 and 180 degree projections
! with the reference's interpolation when no exact angle exists (:140-228),
! the centre-of-rotation / tilt search (:231-364), the 3x3x3 median filter
! with clamped neighbours (:384-447), the correction (:449-464) and the
! slice -> file/offset map of sinograms{js}.dat (:469-487).  Replaced: the MPI
! master/worker projection queue by one GPU pass over all projections.
program projections_to_sinogram_hip
  use iso_c_binding
  use acumen_io
  implicit none
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
    function acu_prep_find_center(c, n, p0, p180, theta, shift, nev) bind(C, name="acu_prep_find_center") result(rc)
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: c
      integer(c_int), value :: n
      real(c_float), intent(in) :: p0(*), p180(*)
      real(c_float), intent(out) :: theta, shift
      integer(c_int), intent(out) :: nev
      integer(c_int) :: rc
    end function
    function acu_prep_sinograms(c, n, nfiles, proj, theta, shift, sino) bind(C, name="acu_prep_sinograms") result(rc)
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: c
      integer(c_int), value :: n, nfiles
      real(c_float), intent(in) :: proj(*)
      real(c_float), value :: theta, shift
      real(c_float), intent(out) :: sino(*)
      integer(c_int) :: rc
    end function
  end interface
  integer :: nfiles, nimage, n_files_sino, n_files_rec, add_Nx, nw, job, j, jy, nev
  integer(kind=8) :: offset, nn
  real :: rtmp, rtmp2, best_theta, best_shift
  real(kind=8) :: t0
  real, allocatable :: f_theta(:), proj(:,:,:), sino(:,:,:), p0(:,:), p180(:,:)
  type(c_ptr) :: ctx
  character(len=150) :: root_dir, exp_name
  character(len=:), allocatable :: rdir, info

  call get_environment_variable("ROOT_CT_DIR", root_dir)
  call get_environment_variable("EXP_CT_NAME", exp_name)
  rdir = rec_dir(root_dir, exp_name)
  info = rdir // "fileinfo_projections.dat"
  t0 = wall_seconds()
  rtmp = 0.0
  call read_real_at(info, 0_8, rtmp); nfiles = int(rtmp)
  call read_real_at(info, 1_8, rtmp); nimage = int(rtmp)
  call read_real_at(info, 2_8, rtmp); n_files_sino = int(rtmp)
  call read_real_at(info, 3_8, rtmp); n_files_rec = int(rtmp)
  write(*,*) "Working on: ", trim(exp_name), " With number of projections: ", nfiles
  allocate(f_theta(nfiles))
  call read_reals_at(rdir // "theta.dat", 0_8, nfiles, f_theta)
  add_Nx = 0
  if (mod(nimage, 2) == 0) add_Nx = 1
  nw = nimage + add_Nx
  nn = int(nimage, 8) * int(nimage, 8)

  ! every projection, from the file and offset of the reference's split
  allocate(proj(nimage, nimage, nfiles))
  rtmp = real(nfiles) / real(n_files_sino)
  do job = 1, nfiles
    j = ceiling(job / rtmp)
    offset = int(job - floor((j - 1) * rtmp) - 1, 8) * nn
    call read_reals_at(rdir // numbered("projections", j), offset, int(nn), proj(:, :, job))
  end do

  allocate(p0(nimage, nimage), p180(nimage, nimage))
  call projection_at(0.0, p0)
  call projection_at(180.0, p180)
  if (acu_ctx_create(0_c_int, ctx) /= 0) call fail("acu_ctx_create")
  if (acu_prep_find_center(ctx, nimage, p0, p180, best_theta, best_shift, nev) /= 0) &
    call fail("acu_prep_find_center")
  write(*,*) "Best theta: ", best_theta, " Best shift: ", best_shift

  allocate(sino(nw, nfiles, nimage))
  if (acu_prep_sinograms(ctx, nimage, nfiles, proj, best_theta, best_shift, sino) /= 0) &
    call fail("acu_prep_sinograms")
  ! slice jy of sinograms{j}.dat: rows of all projections, Nw values each
  rtmp2 = real(nimage) / real(n_files_sino)
  do jy = 1, nimage
    j = ceiling(jy / rtmp2)
    offset = int(jy - floor((j - 1) * rtmp2) - 1, 8) * int(nw, 8) * int(nfiles, 8)
    call write_reals_at(rdir // numbered("sinograms", j), offset, nw * nfiles, sino(:, :, jy))
  end do
  if (acu_ctx_destroy(ctx) /= 0) call fail("acu_ctx_destroy")
  write(*,*) "Total Compute time: ", wall_seconds() - t0

contains

  ! The projection at `target` degrees: the first closest angle, or the
  ! linear interpolation with its neighbour on the far side (:140-228).
  subroutine projection_at(target, out)
    real, intent(in) :: target
    real, intent(out) :: out(:,:)
    integer :: js
    js = minloc(abs(f_theta - target), DIM=1)
    if (f_theta(js) == target) then
      out = proj(:, :, js)
    else if (f_theta(js) <= target) then
      if (js < nfiles) then
        out = (target - f_theta(js)) * (proj(:, :, js + 1) - proj(:, :, js)) / (f_theta(js + 1) - f_theta(js)) &
              + proj(:, :, js)
      else
        out = proj(:, :, js)
      end if
    else
      if (js > 1) then
        out = (target - f_theta(js - 1)) * (proj(:, :, js) - proj(:, :, js - 1)) / (f_theta(js) - f_theta(js - 1)) &
              + proj(:, :, js - 1)
      else
        out = proj(:, :, js)
      end if
    end if
  end subroutine projection_at

  subroutine fail(what)
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    call c_f_pointer(acu_last_error(), msg, [1024])
    k = 0
    do while (k < 1024)
      if (msg(k + 1) == c_null_char) exit
      k = k + 1
    end do
    write(*,*) "acumen: ", what, " failed: ", (msg(j), j = 1, k)
    error stop 2
  end subroutine fail

end program projections_to_sinogram_hip
