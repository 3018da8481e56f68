! tv_hip.F90 -- FBP + split-Bregman TV of every slice of a volume on one GPU:
! the driver of tomography/tv_hpc.F90:392-652 without MPI.
!
! Kept from the reference: the environment contract (ROOT_CT_DIR,
! EXP_CT_NAME, Sub_Sample_Radial, Sub_Sample_Theta, Ninner_Loops,
! Nouter_Loops, Opt_ncg, Opt_lambda, Opt_mu, Radon_wpu), the header read of
! fileinfo_projections.dat offsets 0-3, the output naming (fbp / fbp_tv, or
! fbp_sub / fbp_sub_tv when sub-sampling), the angle window, radial and
! angular sub-sampling, odd padding and every slice -> file/offset map.
! Replaced: the MPI master/worker queue (tv_hpc.F90:583-631) by batches of
! ACUMEN_BATCH slices (default 16) reconstructed together on the GPU.
! Added: a per-slice done map, {fbp_tv|fbp_sub_tv}.done beside the outputs,
! marked as each batch's images are written; ACUMEN_RESUME=1 skips the
! slices an earlier run with the same parameters marked, and
! ACUMEN_MAX_BATCHES=k stops after k batches (a run split over jobs).
program tv_hip
  use Tomography
  use acumen_io
  implicit none
  integer :: nsub, nsub_t, nfiles, n_files_sino, n_files_rec, nimagef, jtheta_s, jtheta_e
  integer :: jx, jy, js, job, b0, nb, i, batch, ios, jouter, k, npend, nbatches, max_batches
  logical :: resume
  integer, allocatable :: todo(:)
  integer(kind=1), allocatable :: done(:)
  integer(kind=8) :: offset
  real :: rtmp, rtmp2, rrec
  real(kind=8) :: t0
  real, allocatable :: f_theta(:), sino(:,:), fslice(:,:), img_fbp(:,:), img_tv(:,:), delta(:)
  character(len=150) :: root_dir, exp_name, c_string
  character(len=:), allocatable :: rdir, info, fbp_name, tv_name, done_path

  call get_environment_variable("ROOT_CT_DIR", root_dir)
  call get_environment_variable("EXP_CT_NAME", exp_name)
  call get_environment_variable("Sub_Sample_Radial", c_string); read(c_string, *) nsub
  call get_environment_variable("Sub_Sample_Theta", c_string); read(c_string, *) nsub_t
  call get_environment_variable("Ninner_Loops", c_string); read(c_string, *) ninnner
  call get_environment_variable("Nouter_Loops", c_string); read(c_string, *) nouter
  call get_environment_variable("Opt_ncg", c_string); read(c_string, *) ncg
  call get_environment_variable("Opt_lambda", c_string); read(c_string, *) lambda
  call get_environment_variable("Opt_mu", c_string); read(c_string, *) mu
  call get_environment_variable("Radon_wpu", c_string); read(c_string, *) wpu
  call get_environment_variable("ACUMEN_BATCH", c_string)
  batch = 16
  if (len_trim(c_string) > 0) then
    read(c_string, *, iostat=ios) batch
    if (ios /= 0 .or. batch < 1) batch = 16
  end if
  call get_environment_variable("ACUMEN_RESUME", c_string)
  resume = trim(c_string) == "1"
  call get_environment_variable("ACUMEN_MAX_BATCHES", c_string)
  max_batches = 0
  if (len_trim(c_string) > 0) then
    read(c_string, *, iostat=ios) max_batches
    if (ios /= 0 .or. max_batches < 0) max_batches = 0
  end if
  write(*,*) "Running: ", trim(exp_name)
  write(*,*) " Itr: ", nouter
  write(*,*) " Lambda: ", lambda
  write(*,*) " Mu: ", mu
  t0 = wall_seconds()

  rdir = rec_dir(root_dir, exp_name)
  info = rdir // "fileinfo_projections.dat"
  rtmp = 0.0
  call read_real_at(info, 0_8, rtmp); nfiles = int(rtmp)
  call read_real_at(info, 1_8, rtmp); nimagef = int(rtmp)
  call read_real_at(info, 2_8, rtmp); n_files_sino = int(rtmp)
  call read_real_at(info, 3_8, rtmp); n_files_rec = int(rtmp)

  if (nsub + nsub_t .EQ. 2) then
    fbp_name = "fbp"
    tv_name = "fbp_tv"
  else
    fbp_name = "fbp_sub"
    tv_name = "fbp_sub_tv"
  end if

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
  if ((nimage - 1) * nsub .NE. nimagef + add_Nx - 1) write(*,*) "Warning:  nsub must evenly sub-divide"
  write(*,*) "Reconstruction with Resolution: ", nimage, "with", ntheta, "projections"

  allocate(sino(nimagef + add_Nx, nfiles), fslice(nimage, ntheta))
  allocate(img_fbp(nimage, nimage), img_tv(nimage, nimage), delta(max(nouter, 1)))
  rtmp2 = real(nimagef) / real(n_files_sino)
  rrec = real(nimagef) / real(n_files_rec)
  allocate(done(nimagef))
  done_path = rdir // tv_name // ".done"
  call done_map_open(done_path, [nimagef, nimage, ntheta, nsub, nsub_t, wpu, ninnner, nouter, ncg, 0], &
                     [mu, lambda], resume, done)
  todo = pack([(k, k = 1, nimagef)], done == 0)
  npend = size(todo)
  if (npend < nimagef) write(*,*) "Resuming:", nimagef - npend, "slices done earlier"
  if (npend > 0) call batch_init(min(batch, npend))
  nbatches = 0
  do b0 = 1, npend, batch
    if (max_batches > 0 .and. nbatches >= max_batches) then
      write(*,*) "Stopped after", nbatches, "batches;", npend - b0 + 1, "slices pending (ACUMEN_RESUME=1 continues)"
      exit
    end if
    nb = min(batch, npend - b0 + 1)
    if (nb /= min(batch, npend)) call batch_init(nb)
    do i = 1, nb
      job = todo(b0 + i - 1)
      js = ceiling(job / rtmp2)
      offset = int(job - floor((js - 1) * rtmp2) - 1, 8) * int(nimagef + add_Nx, 8) * int(nfiles, 8)
      call read_reals_at(rdir // numbered("sinograms", js), offset, (nimagef + add_Nx) * nfiles, sino)
      do jy = 1, ntheta
        do jx = 1, nimage
          fslice(jx, jy) = sino((jx - 1) * nsub + 1, (jy - 1) * nsub_t + jtheta_s)
        end do
      end do
      call batch_set_slice(i, fslice)
    end do
    call batch_tvopt()
    do i = 1, nb
      job = todo(b0 + i - 1)
      call batch_get(i, ACU_IMG_FBP, img_fbp)
      call batch_get(i, ACU_IMG_TV, img_tv)
      if (nouter > 0) then
        call batch_delta(i, delta)
        do jouter = 50, nouter, 50
          write(*,*) "Slice:", job, "Loop: ", jouter, "delta:", delta(jouter)
        end do
      end if
      js = ceiling(job / rrec)
      offset = int(job - floor((js - 1) * rrec) - 1, 8) * int(nimage, 8) * int(nimage, 8)
      call write_reals_at(rdir // numbered(fbp_name, js), offset, nimage * nimage, img_fbp)
      call write_reals_at(rdir // numbered(tv_name, js), offset, nimage * nimage, img_tv)
      call done_map_mark(done_path, job)
    end do
    nbatches = nbatches + 1
    write(*,*) "Slices", todo(b0), "to", todo(b0 + nb - 1), "done"
  end do
  call fin_radon()
  write(*,*) "Total Compute time: ", real(wall_seconds() - t0)
end program tv_hip
