! diablo_gpu.f90 -- iso_c_binding interface to libdiablo_gpu.so (include/diablo_gpu.h).
!
! This is the module a DIABLO build adds to bind the MI355X time step.  The
! derived type dgpu_config is interoperable with the C struct of the same name:
! same members in the same order, so the companion C compiler lays both out alike.
! Array arguments are passed with c_loc() of the reference's own arrays, e.g.
!     complex(rkind) :: cu1(0:Nxp, 0:Nz+1, 0:Nyp+1)        (for/fft.f90:71-96)
module diablo_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: DGPU_FP64 = 0, DGPU_MIXED = 1
  integer(c_int32_t), parameter, public :: DGPU_BC_DIRICHLET = 0, DGPU_BC_NEUMANN = 1
  integer, parameter, public :: DGPU_MAX_TH = 4

  type, bind(C), public :: dgpu_config
    integer(c_int32_t) :: nx, ny, nz, n_th
    integer(c_int32_t) :: precision, use_les, les_model_type, flavor_front, f_type
    integer(c_int32_t) :: jstart, jend
    integer(c_int32_t) :: jstart_th(DGPU_MAX_TH), jend_th(DGPU_MAX_TH)
    integer(c_int32_t) :: u_bc_ymin, v_bc_ymin, w_bc_ymin
    integer(c_int32_t) :: u_bc_ymax, v_bc_ymax, w_bc_ymax
    integer(c_int32_t) :: th_bc_ymin(DGPU_MAX_TH), th_bc_ymax(DGPU_MAX_TH)
    integer(c_int32_t) :: homogeneous_x
    integer(c_int32_t) :: reserved_i(3)
    real(c_double) :: lx, ly, lz
    real(c_double) :: nu, beta, nu_v_scale
    real(c_double) :: ro_inv, delta
    real(c_double) :: grav_x, grav_y, grav_z
    real(c_double) :: px0, omega0, amp_omega0, force_start
    real(c_double) :: u_bc_ymin_c1, v_bc_ymin_c1, w_bc_ymin_c1
    real(c_double) :: u_bc_ymax_c1, v_bc_ymax_c1, w_bc_ymax_c1
    real(c_double) :: th_bc_ymin_c1(DGPU_MAX_TH), th_bc_ymax_c1(DGPU_MAX_TH)
    real(c_double) :: pr(DGPU_MAX_TH), dthdx(DGPU_MAX_TH), dthdz(DGPU_MAX_TH)
    real(c_double) :: reserved_d(4)
    type(c_ptr) :: gy, gyf, dy, dyf
    type(c_ptr) :: kx, kz
  end type dgpu_config

  ! the Timestep attributes of a flow file (include/dgpu_io.h)
  type, bind(C), public :: dgpu_flow_times
    real(c_double) :: time, save_flow_time, save_stats_time, save_movie_time
    integer(c_int32_t) :: time_step
  end type dgpu_flow_times

  public :: dgpu_create, dgpu_destroy, dgpu_upload_state, dgpu_download_state
  public :: dgpu_upload_rk, dgpu_download_rk, dgpu_rk_substep, dgpu_step, dgpu_rem_div
  public :: dgpu_fft_xz_to_fourier, dgpu_fft_xz_to_physical, dgpu_synchronize
  public :: dgpu_last_step_ms, dgpu_field_doubles, dgpu_last_error_f
  public :: dgpu_download_les, dgpu_stats, dgpu_stats_doubles
  public :: dgpu_set_variable_dt, dgpu_courant, dgpu_set_closure, dgpu_filter_chan
  public :: dgpu_write_flow, dgpu_read_flow, dgpu_poisson_p, dgpu_stats_movie, dgpu_stats_movie_sizes

  interface
    integer(c_int) function dgpu_create(cfg, h) bind(C, name="dgpu_create")
      import :: c_int, c_ptr, dgpu_config
      type(dgpu_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function dgpu_destroy(h) bind(C, name="dgpu_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dgpu_upload_state(h, cu1, cu2, cu3, cp, cth) bind(C, name="dgpu_upload_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, cu1, cu2, cu3, cp, cth
    end function
    integer(c_int) function dgpu_download_state(h, cu1, cu2, cu3, cp, cth) bind(C, name="dgpu_download_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, cu1, cu2, cu3, cp, cth
    end function
    integer(c_int) function dgpu_upload_rk(h, cf1, cf2, cf3, cfth) bind(C, name="dgpu_upload_rk")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, cf1, cf2, cf3, cfth
    end function
    integer(c_int) function dgpu_download_rk(h, cf1, cf2, cf3, cfth) bind(C, name="dgpu_download_rk")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, cf1, cf2, cf3, cfth
    end function
    integer(c_int) function dgpu_rk_substep(h, rk_step, delta_t, time, time_step, create_new_flow) &
        bind(C, name="dgpu_rk_substep")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: rk_step, time_step, create_new_flow
      real(c_double), value :: delta_t, time
    end function
    integer(c_int) function dgpu_step(h, delta_t, time, time_step, create_new_flow) bind(C, name="dgpu_step")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: h
      real(c_double), value :: delta_t, time
      integer(c_int32_t), value :: time_step, create_new_flow
    end function
    integer(c_int) function dgpu_rem_div(h) bind(C, name="dgpu_rem_div")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dgpu_poisson_p(h) bind(C, name="dgpu_poisson_p")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dgpu_fft_xz_to_fourier(h, phys, four) bind(C, name="dgpu_fft_xz_to_fourier")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, phys, four
    end function
    integer(c_int) function dgpu_fft_xz_to_physical(h, four, phys) bind(C, name="dgpu_fft_xz_to_physical")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, four, phys
    end function
    integer(c_int) function dgpu_synchronize(h) bind(C, name="dgpu_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dgpu_last_step_ms(h, ms) bind(C, name="dgpu_last_step_ms")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: ms
    end function
    integer(c_int64_t) function dgpu_field_doubles(cfg) bind(C, name="dgpu_field_doubles")
      import :: c_int64_t, dgpu_config
      type(dgpu_config), intent(in) :: cfg
    end function
    integer(c_int) function dgpu_download_les(h, nu_t, kappa_t) bind(C, name="dgpu_download_les")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, nu_t, kappa_t
    end function
    integer(c_int) function dgpu_stats(h, time, out) bind(C, name="dgpu_stats")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: time
      type(c_ptr), value :: out
    end function
    integer(c_int64_t) function dgpu_stats_doubles(cfg) bind(C, name="dgpu_stats_doubles")
      import :: c_int64_t, dgpu_config
      type(dgpu_config), intent(in) :: cfg
    end function
    integer(c_int) function dgpu_stats_movie(h, time, nxm, nym, nzm, planes, slices) &
        bind(C, name="dgpu_stats_movie")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: time
      integer(c_int32_t), value :: nxm, nym, nzm
      type(c_ptr), value :: planes, slices
    end function
    integer(c_int) function dgpu_stats_movie_sizes(cfg, n_planes, n_slices) bind(C, name="dgpu_stats_movie_sizes")
      import :: c_int, c_int64_t, dgpu_config
      type(dgpu_config), intent(in) :: cfg
      integer(c_int64_t), intent(out) :: n_planes, n_slices
    end function
    integer(c_int) function dgpu_set_variable_dt(h, variable_dt, update_dt, cfl) bind(C, name="dgpu_set_variable_dt")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: variable_dt, update_dt
      real(c_double), value :: cfl
    end function
    integer(c_int) function dgpu_courant(h, time, time_limit, save_stats_time, save_flow_time, delta_t, ran) &
        bind(C, name="dgpu_courant")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: time, time_limit, save_stats_time, save_flow_time
      real(c_double), intent(out) :: delta_t
      integer(c_int32_t), intent(out) :: ran
    end function
    integer(c_int) function dgpu_filter_chan(h, n) bind(C, name="dgpu_filter_chan")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: n
    end function
    ! les_model_type 6 (the learned closure, include/dgpu_closure.h): parameters
    ! and normaliser as written by diablo3_amd.train_closure
    integer(c_int) function dgpu_set_closure(h, params, norm) bind(C, name="dgpu_set_closure")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, params, norm
    end function
    ! flow checkpoints in the reference's HDF5 layout (libdgpu_io.so): they replace
    ! WriteHDF5 / ReadHDF5 (for/phdf5.f90:918-1456) in save_flow / read_flow;
    ! fname is passed as trim(fname)//c_null_char
    integer(c_int) function dgpu_write_flow(h, fname, t, save_pressure) bind(C, name="dgpu_write_flow")
      import :: c_int, c_int32_t, c_ptr, c_char, dgpu_flow_times
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*), intent(in) :: fname
      type(dgpu_flow_times), intent(in) :: t
      integer(c_int32_t), value :: save_pressure
    end function
    integer(c_int) function dgpu_read_flow(h, fname, th_mask, t, read_pressure) bind(C, name="dgpu_read_flow")
      import :: c_int, c_int32_t, c_ptr, c_char, dgpu_flow_times
      type(c_ptr), value :: h
      character(kind=c_char), dimension(*), intent(in) :: fname
      integer(c_int32_t), value :: th_mask  ! bit n - 1: read TH n (not create_new_th(n))
      type(dgpu_flow_times), intent(out) :: t
      integer(c_int32_t), intent(out) :: read_pressure
    end function
    type(c_ptr) function dgpu_last_error_c() bind(C, name="dgpu_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! dgpu_last_error as a Fortran string
  function dgpu_last_error_f() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    interface
      integer(c_size_t) function strlen(s) bind(C, name="strlen")
        import :: c_size_t, c_ptr
        type(c_ptr), value :: s
      end function
    end interface
    type(c_ptr) :: cp
    cp = dgpu_last_error_c()
    n = int(strlen(cp))
    call c_f_pointer(cp, p, [n])
    allocate (character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function

end module diablo_gpu
