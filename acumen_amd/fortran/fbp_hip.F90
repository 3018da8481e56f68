! fbp_hip.F90 -- filtered backprojection of every slice of a volume on one
! GPU: the driver of tomography/fbp_hpc.F90:156-424 without MPI.
!
! Kept from the reference: environment ROOT_CT_DIR / EXP_CT_NAME, the
! input_parm.txt read and its echo into fileinfo_projections.dat offsets 2-9,
! the header read of offsets 0-11 (file counts from offsets 10-11), the angle
! window, odd padding, the slice -> file/offset maps of sinograms{js}.dat and
! fbp{js}.dat, and the FP32 layouts.  Replaced: the MPI master/worker slice
! queue (fbp_hpc.F90:357-403) by batches of ACUMEN_BATCH slices (default 16)
! reconstructed together on the GPU.
program fbp_hip
  use Tomography
  use acumen_io
  implicit none
  integer :: nsub, nsub_t, nfiles, n_files_sino, n_files_rec, nimagef, jtheta_s, jtheta_e
  integer :: jx, jy, js, job, b0, nb, i, batch, ios
  integer(kind=8) :: offset
  real :: rtmp, rtmp2, rrec
  real(kind=8) :: t0
  real, allocatable :: f_theta(:), sino(:,:), fslice(:,:), img(:,:)
  character(len=150) :: root_dir, exp_name, c_string
  character(len=:), allocatable :: rdir, info

  call get_environment_variable("ROOT_CT_DIR", root_dir)
  call get_environment_variable("EXP_CT_NAME", exp_name)
  call get_environment_variable("ACUMEN_BATCH", c_string)
  batch = 16
  if (len_trim(c_string) > 0) then
    read(c_string, *, iostat=ios) batch
    if (ios /= 0 .or. batch < 1) batch = 16
  end if
  rdir = rec_dir(root_dir, exp_name)
  info = rdir // "fileinfo_projections.dat"
  t0 = wall_seconds()

  write(*,*) 'Updating from file: ', trim(root_dir) // "/" // trim(exp_name) // "/input_parm.txt"
  call read_input_parm(trim(root_dir) // "/" // trim(exp_name) // "/input_parm.txt", &
                       nsub, nsub_t, ninnner, nouter, ncg, lambda, mu, wpu)
  ! echo of the parameters into the header (fbp_hpc.F90:220-235)
  call write_real_at(info, 2_8, real(nsub))
  call write_real_at(info, 3_8, real(nsub_t))
  call write_real_at(info, 4_8, real(ninnner))
  call write_real_at(info, 5_8, real(nouter))
  call write_real_at(info, 6_8, real(ncg))
  call write_real_at(info, 7_8, lambda)
  call write_real_at(info, 8_8, mu)
  call write_real_at(info, 9_8, real(wpu))
  ! header read (fbp_hpc.F90:240-275)
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

  ! full resolution for FBP (fbp_hpc.F90:279-281)
  nsub = 1
  nsub_t = 1

  allocate(f_theta(nfiles))
  call read_reals_at(rdir // "theta.dat", 0_8, nfiles, f_theta)
  call angle_window(f_theta, jtheta_s, jtheta_e)
  ntheta = jtheta_e - jtheta_s + 1
  add_Nx = 0
  if (mod(nimagef, 2) .EQ. 0) add_Nx = 1
  nimage = nimagef + add_Nx

  call int_radon()
  do jy = 1, ntheta
    theta(jy) = f_theta((jy - 1) + jtheta_s)
  end do
  write(*,*) "Reconstruction with Resolution: ", nimage, "with", ntheta, "projections"

  allocate(sino(nimagef + add_Nx, nfiles), fslice(nimage, ntheta), img(nimage, nimage))
  rtmp2 = real(nimagef) / real(n_files_sino)
  rrec = real(nimagef) / real(n_files_rec)
  call batch_init(min(batch, nimagef))
  do b0 = 1, nimagef, batch
    nb = min(batch, nimagef - b0 + 1)
    if (nb /= min(batch, nimagef)) call batch_init(nb)
    do i = 1, nb
      job = b0 + i - 1
      js = ceiling(job / rtmp2)
      offset = int(job - floor((js - 1) * rtmp2) - 1, 8) * int(nimagef + add_Nx, 8) * int(nfiles, 8)
      call read_reals_at(rdir // numbered("sinograms", js), offset, (nimagef + add_Nx) * nfiles, sino)
      do jy = 1, ntheta
        do jx = 1, nimage
          fslice(jx, jy) = sino(jx, (jy - 1) + jtheta_s)
        end do
      end do
      call batch_set_slice(i, fslice)
    end do
    call batch_fbp()
    do i = 1, nb
      job = b0 + i - 1
      call batch_get(i, ACU_IMG_FBP, img)
      js = ceiling(job / rrec)
      offset = int(job - floor((js - 1) * rrec) - 1, 8) * int(nimage, 8) * int(nimage, 8)
      call write_reals_at(rdir // numbered("fbp", js), offset, nimage * nimage, img)
    end do
    write(*,*) "Slices", b0, "to", b0 + nb - 1, "done"
  end do
  call fin_radon()
  write(*,*) "Total Compute time: ", real(wall_seconds() - t0)
end program fbp_hip
