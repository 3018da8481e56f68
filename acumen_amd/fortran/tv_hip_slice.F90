! tv_hip_slice.F90 -- FBP + split-Bregman TV of the middle slice on one GPU:
! the driver of tomography/tv_hpc_slice.F90:368-588 (which already runs on
! one rank).
!
! Kept from the reference: ROOT_CT_DIR / EXP_CT_NAME, input_parm.txt, the
! echo of its values into fileinfo_projections.dat offsets 2-9 and the header
! read of offsets 0-11, the angle window and sub-sampling, the middle slice
! nimage/2 read from sinograms1.dat, and the outputs fbp_tv_slice.dat and
! fbp_sub_slice.dat.
program tv_hip_slice
  use Tomography
  use acumen_io
  implicit none
  integer :: nsub, nsub_t, nfiles, n_files_sino, n_files_rec, nimagef, jtheta_s, jtheta_e
  integer :: jx, jy, job_id
  integer(kind=8) :: offset
  real :: rtmp
  real(kind=8) :: t0
  real, allocatable :: f_theta(:), sino(:,:), img_fbp(:,:), img_tv(:,:)
  character(len=150) :: root_dir, exp_name
  character(len=:), allocatable :: rdir, info

  call get_environment_variable("ROOT_CT_DIR", root_dir)
  call get_environment_variable("EXP_CT_NAME", exp_name)
  rdir = rec_dir(root_dir, exp_name)
  info = rdir // "fileinfo_projections.dat"
  t0 = wall_seconds()

  call read_input_parm(trim(root_dir) // "/" // trim(exp_name) // "/input_parm.txt", &
                       nsub, nsub_t, ninnner, nouter, ncg, lambda, mu, wpu)
  call write_real_at(info, 2_8, real(nsub))
  call write_real_at(info, 3_8, real(nsub_t))
  call write_real_at(info, 4_8, real(ninnner))
  call write_real_at(info, 5_8, real(nouter))
  call write_real_at(info, 6_8, real(ncg))
  call write_real_at(info, 7_8, lambda)
  call write_real_at(info, 8_8, mu)
  call write_real_at(info, 9_8, real(wpu))
  rtmp = 0.0
  call read_real_at(info, 0_8, rtmp); nfiles = int(rtmp)
  call read_real_at(info, 1_8, rtmp); nimagef = int(rtmp)
  call read_real_at(info, 2_8, rtmp); nsub = int(rtmp)
  call read_real_at(info, 3_8, rtmp); nsub_t = int(rtmp)
  call read_real_at(info, 4_8, rtmp); ninnner = int(rtmp)
  call read_real_at(info, 5_8, rtmp); nouter = int(rtmp)
  call read_real_at(info, 6_8, rtmp); ncg = int(rtmp)
  call read_real_at(info, 7_8, lambda)
  call read_real_at(info, 8_8, mu)
  call read_real_at(info, 9_8, rtmp); wpu = int(rtmp)
  call read_real_at(info, 10_8, rtmp); n_files_sino = int(rtmp)
  call read_real_at(info, 11_8, rtmp); n_files_rec = int(rtmp)

  allocate(f_theta(nfiles))
  call read_reals_at(rdir // "theta.dat", 0_8, nfiles, f_theta)
  call angle_window(f_theta, jtheta_s, jtheta_e)
  ntheta = floor(real(jtheta_e - jtheta_s + 1) / real(nsub_t))
  add_Nx = 0
  if (mod(nimagef, 2) .EQ. 0) add_Nx = 1
  nimage = floor(real(nimagef + add_Nx - 1) / real(nsub)) + 1

  call int_radon()
  do jy = 1, ntheta
    theta(jy) = f_theta((jy - 1) * nsub_t + jtheta_s)
  end do

  job_id = nimage / 2
  allocate(sino(nimagef + add_Nx, nfiles), img_tv(nimage, nimage), img_fbp(nimage, nimage))
  offset = int(job_id - 1, 8) * int(nimagef + add_Nx, 8) * int(nfiles, 8)
  call read_reals_at(rdir // "sinograms1.dat", offset, (nimagef + add_Nx) * nfiles, sino)
  do jy = 1, ntheta
    do jx = 1, nimage
      f(jx, jy) = sino((jx - 1) * nsub + 1, (jy - 1) * nsub_t + jtheta_s)
    end do
  end do

  call tvopt(img_tv, img_fbp, job_id)

  call write_reals_at(rdir // "fbp_tv_slice.dat", 0_8, nimage * nimage, img_tv)
  call write_reals_at(rdir // "fbp_sub_slice.dat", 0_8, nimage * nimage, img_fbp)
  call fin_radon()
  write(*,*) "Total Compute time: ", real(wall_seconds() - t0)
end program tv_hip_slice
