! dgpu_drive.f90 -- a Fortran host driving libdiablo_gpu.so through iso_c_binding.
!
! It declares the state exactly as DIABLO does (complex(8) cu(0:Nxp,0:Nz+1,0:Nyp+1),
! for/fft.f90:71-96), uploads it, runs the reference's time loop
!     do rk_step = 1,3 : call rk_chan_1          (for/diablo.f90:62-72)
! with dgpu_rk_substep in place of rk_chan_1, and downloads the result.
!
! Usage: dgpu_drive <input file> <output file>
! Input (stream, native endianness): int32 magic 1146114389, int32 nsteps,
! real(8) delta_t, int32 create_new_flow, the raw dgpu_config, gy gyf dy dyf
! (Ny+2 each), kx (Nx/2), kz (2(Nz/3)+1), then cu1 cu2 cu3 cp cth(1:N_th).
! Output: cu1 cu2 cu3 cp cth(1:N_th) after nsteps steps.
program dgpu_drive
  use, intrinsic :: iso_c_binding
  use diablo_gpu
  implicit none
  type(dgpu_config), target :: cfg
  type(c_ptr) :: h
  real(c_double), allocatable, target :: gy(:), gyf(:), dy(:), dyf(:), kx(:), kz(:)
  complex(c_double_complex), allocatable, target :: cu1(:,:,:), cu2(:,:,:), cu3(:,:,:), cp(:,:,:)
  complex(c_double_complex), allocatable, target :: cth(:,:,:,:)
  integer(c_int32_t) :: magic, nsteps, create_new_flow, time_step, rk_step
  real(c_double) :: delta_t, time
  integer :: nxp, nz, nyp, nth, rc, n
  integer(c_int8_t), allocatable :: raw(:)
  character(len=4096) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (10, file=trim(fin), access='stream', form='unformatted', status='old')
  read (10) magic, nsteps, delta_t, create_new_flow
  if (magic /= 1146114389) error stop 'dgpu_drive: bad input file'
  ! the struct arrives as its raw C bytes (its pointer members are reset below)
  allocate (raw(c_sizeof(cfg)))
  read (10) raw
  cfg = transfer(raw, cfg)
  nxp = cfg%nx / 2
  nz = cfg%nz
  nyp = cfg%ny
  nth = max(cfg%n_th, 1)
  allocate (gy(0:nyp+1), gyf(0:nyp+1), dy(0:nyp+1), dyf(0:nyp+1))
  allocate (kx(0:nxp-1), kz(0:2*(nz/3)))
  read (10) gy, gyf, dy, dyf, kx, kz
  allocate (cu1(0:nxp, 0:nz+1, 0:nyp+1), cu2(0:nxp, 0:nz+1, 0:nyp+1))
  allocate (cu3(0:nxp, 0:nz+1, 0:nyp+1), cp(0:nxp, 0:nz+1, 0:nyp+1))
  allocate (cth(0:nxp, 0:nz+1, 0:nyp+1, 1:nth))
  cth = (0.d0, 0.d0)
  read (10) cu1, cu2, cu3, cp
  do n = 1, cfg%n_th
    read (10) cth(:, :, :, n)
  end do
  close (10)
  cfg%gy = c_loc(gy)
  cfg%gyf = c_loc(gyf)
  cfg%dy = c_loc(dy)
  cfg%dyf = c_loc(dyf)
  cfg%kx = c_loc(kx)
  cfg%kz = c_loc(kz)

  rc = dgpu_create(cfg, h)
  if (rc /= 0) then
    write (*, *) 'dgpu_create: ', dgpu_last_error_f()
    error stop 1
  end if
  rc = dgpu_upload_state(h, c_loc(cu1), c_loc(cu2), c_loc(cu3), c_loc(cp), c_loc(cth))
  if (rc /= 0) error stop 'dgpu_upload_state failed'
  time = 0.d0
  do time_step = 1, nsteps
    do rk_step = 1, 3
      rc = dgpu_rk_substep(h, rk_step, delta_t, time, time_step, create_new_flow)
      if (rc /= 0) then
        write (*, *) 'dgpu_rk_substep: ', dgpu_last_error_f()
        error stop 1
      end if
    end do
    time = time + delta_t
  end do
  rc = dgpu_download_state(h, c_loc(cu1), c_loc(cu2), c_loc(cu3), c_loc(cp), c_loc(cth))
  if (rc /= 0) error stop 'dgpu_download_state failed'
  rc = dgpu_destroy(h)

  open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (11) cu1, cu2, cu3, cp
  do n = 1, cfg%n_th
    write (11) cth(:, :, :, n)
  end do
  close (11)
  write (*, '(A, I0, A)') 'dgpu_drive: ', nsteps, ' steps done'
end program dgpu_drive
