! tandem_hip.f90 -- iso_c_binding interface of libtandem_hip.so for a Fortran host.
!
! The reference's hot path lives in module procedures of subroutines3d
! (src/Subroutines3D.f90: mpigo :16, deriv :84, filte :145) and inline in the
! time loop of src/Main3D.f90:552-965.  This module declares the C entry
! points of include/tandem_hip.h so that the Fortran host can keep its call
! structure and replace those calls one by one (per-call drop-ins) or the
! whole time loop (tp_step).  Every function returns 0 on success; tp_check
! stops the program with the library's message otherwise, which is how the
! reference reacts to a failure (MPI default abort handler / error stop).
module tandem_hip
  use iso_c_binding, only: c_ptr, c_null_ptr, c_int32_t, c_int64_t, c_double, c_char, &
                           c_associated, c_f_pointer
  implicit none
  private :: c_strlen

  ! return codes
  integer(c_int32_t), parameter :: TP_OK = 0, TP_E_ARG = 1, TP_E_STATE = 2, TP_E_HIP = 3, &
                                   TP_E_SIZE = 4, TP_E_LAYOUT = 5
  ! integer case parameters, iparams(1:16) (0-based slot + 1)
  integer, parameter :: TP_I_NKRK = 1, TP_I_NVISCOUS = 2, TP_I_NSMF = 3, TP_I_LES = 4, &
                        TP_I_NEXTRABC = 5, TP_I_NEXTGCIC = 6, TP_I_INFLOW = 7, TP_I_NMODES = 8, &
                        TP_I_SEED = 9, TP_I_NTZ = 10, TP_I_FORCING = 11, TP_I_NPARAM = 16
  ! double case parameters, dparams(1:48)
  integer, parameter :: TP_D_REOO = 1, TP_D_TEMPOO = 2, TP_D_AMACH1 = 3, TP_D_AMACH2 = 4, &
                        TP_D_AMACH3 = 5, TP_D_WTEMP = 6, TP_D_CFL = 7, TP_D_TMAX = 8, &
                        TP_D_TIMF = 9, TP_D_FLTK = 10, TP_D_FLTKBC = 11, TP_D_DTO = 12, &
                        TP_D_AOA0 = 13, TP_D_AOA1 = 14, TP_D_TPS = 15, TP_D_TP = 16, &
                        TP_D_SMAGO1 = 17, TP_D_SMAGO2 = 18, TP_D_DOMLEN = 19, TP_D_SZTH1 = 20, &
                        TP_D_SZTH2 = 21, TP_D_SZXT = 22, TP_D_SZCO = 23, TP_D_SPAN = 24, &
                        TP_D_WLEW = 25, TP_D_WLEA = 26, TP_D_TGUSTO = 27, TP_D_SLIT = 28, &
                        TP_D_TU = 29, TP_D_LTURB = 30, TP_D_GUSTW = 31, TP_D_GUSTK = 32, &
                        TP_D_CFIT = 33, TP_D_AMFOR = 34, TP_D_XFOR = 35, TP_D_YFOR = 36, &
                        TP_D_RFOR = 37, TP_D_TSFOR = 38, TP_D_TEFOR = 39, TP_D_NPARAM = 48
  ! array ids of tp_upload / tp_download (names of src/MainVar3D.f90:78-102)
  integer(c_int32_t), parameter :: TP_QA = 0, TP_QO = 1, TP_DE = 2, TP_RR = 3, TP_SS = 4, &
                                   TP_XIM = 5, TP_ETM = 6, TP_ZEM = 7, TP_P = 8, TP_YACO = 9, &
                                   TP_TV = 10, TP_DRVA1 = 11, TP_DRVA2 = 12, TP_DRVA3 = 13, &
                                   TP_DRVB1 = 14, TP_DRVB2 = 15, TP_DRVB3 = 16, TP_SEND1 = 17, &
                                   TP_SEND2 = 18, TP_SEND3 = 19, TP_RECV1 = 20, TP_RECV2 = 21, &
                                   TP_RECV3 = 22, TP_CM1 = 23, TP_CM2 = 24, TP_CM3 = 25, &
                                   TP_XU = 26, TP_XL = 27, TP_YU = 28, TP_YL = 29, TP_RPEX = 30
  ! scalars of a case, s(1:16) (src/MainVar3D.f90:139-141)
  integer, parameter :: TP_S_N = 1, TP_S_NDT = 2, TP_S_DT = 3, TP_S_DTS = 4, TP_S_DTE = 5, &
                        TP_S_TIMO = 6, TP_S_UMF1 = 7, TP_S_UMF2 = 8, TP_S_UMF3 = 9, &
                        TP_S_DUDT1 = 10, TP_S_DUDT2 = 11, TP_S_DUDT3 = 12, TP_S_DTK = 13, &
                        TP_S_DTKO = 14, TP_S_AOA = 15, TP_S_NSCALAR = 16
  ! options
  integer(c_int32_t), parameter :: TP_OPT_EXACT_LINES = 0, TP_OPT_CHUNKS = 1, TP_OPT_STRESSES = 2

  interface
    integer(c_int32_t) function tp_create(ctx) bind(C, name="tp_create")
      import :: c_ptr, c_int32_t
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int32_t) function tp_destroy(ctx) bind(C, name="tp_destroy")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function tp_last_error() bind(C, name="tp_last_error")
      import :: c_ptr
    end function
    integer(c_int32_t) function tp_add_case(ctx, nblocks, dims, nbc, ncd, nsz, njct, jct, jvel, &
                                            iparams, dparams, case_id) bind(C, name="tp_add_case")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nblocks, njct
      integer(c_int32_t), intent(in) :: dims(*), nbc(*), ncd(*), nsz(*), jct(*), jvel(*), iparams(*)
      real(c_double), intent(in) :: dparams(*)
      integer(c_int32_t), intent(out) :: case_id
    end function
    integer(c_int32_t) function tp_finalize(ctx) bind(C, name="tp_finalize")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function tp_set_grid(ctx, case_id, block, xyz) bind(C, name="tp_set_grid")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, block
      real(c_double), intent(in) :: xyz(*)
    end function
    integer(c_int32_t) function tp_setup_geometry(ctx) bind(C, name="tp_setup_geometry")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function tp_initialo(ctx, case_id) bind(C, name="tp_initialo")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
    end function
    integer(c_int32_t) function tp_upload(ctx, case_id, block, array_id, src, count) &
        bind(C, name="tp_upload")
      import :: c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, block, array_id
      real(c_double), intent(in) :: src(*)
      integer(c_int64_t), value :: count
    end function
    integer(c_int32_t) function tp_download(ctx, case_id, block, array_id, dst, count) &
        bind(C, name="tp_download")
      import :: c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, block, array_id
      real(c_double), intent(out) :: dst(*)
      integer(c_int64_t), value :: count
    end function
    integer(c_int64_t) function tp_array_count(ctx, case_id, block, array_id) &
        bind(C, name="tp_array_count")
      import :: c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, block, array_id
    end function
    integer(c_int32_t) function tp_get_scalars(ctx, case_id, s16) bind(C, name="tp_get_scalars")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      real(c_double), intent(out) :: s16(16)
    end function
    integer(c_int32_t) function tp_set_scalars(ctx, case_id, s16) bind(C, name="tp_set_scalars")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      real(c_double), intent(in) :: s16(16)
    end function
    ! per-call drop-ins for src/Subroutines3D.f90 mpigo/deriv/filte (all sub-blocks)
    integer(c_int32_t) function tp_mpigo(ctx, nt, nrt, n45) bind(C, name="tp_mpigo")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nt, nrt, n45
    end function
    integer(c_int32_t) function tp_deriv(ctx, nn, nz, m) bind(C, name="tp_deriv")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nn, nz, m
    end function
    integer(c_int32_t) function tp_filte(ctx, nn, nz) bind(C, name="tp_filte")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nn, nz
    end function
    ! coarse entry points (src/Main3D.f90:554-561, :567-958, :552-965)
    integer(c_int32_t) function tp_filter_step(ctx) bind(C, name="tp_filter_step")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function tp_stage(ctx, nk) bind(C, name="tp_stage")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
    end function
    ! the sections of a stage (src/Main3D.f90:567-958) as drop-ins, in tp_stage's order
    integer(c_int32_t) function tp_primitives(ctx, nk) bind(C, name="tp_primitives")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
    end function
    integer(c_int32_t) function tp_viscous(ctx) bind(C, name="tp_viscous")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function tp_flux_divergence(ctx, nk) bind(C, name="tp_flux_divergence")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
    end function
    integer(c_int32_t) function tp_sponge(ctx) bind(C, name="tp_sponge")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function tp_gcbc(ctx) bind(C, name="tp_gcbc")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function tp_rk_update(ctx, nk) bind(C, name="tp_rk_update")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nk
    end function
    integer(c_int32_t) function tp_wall(ctx) bind(C, name="tp_wall")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function tp_junction(ctx) bind(C, name="tp_junction")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function tp_interface_avg(ctx) bind(C, name="tp_interface_avg")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    ! deriv on one field in device memory (src, dst: device addresses)
    integer(c_int32_t) function tp_field_deriv(lxi, let, lze, nbc2, nn, src, dst, stream) &
        bind(C, name="tp_field_deriv")
      import :: c_ptr, c_int32_t
      integer(c_int32_t), value :: lxi, let, lze, nn
      integer(c_int32_t), intent(in) :: nbc2(2)
      type(c_ptr), value :: src, dst, stream
    end function
    ! the same without allocating (work area held by the library per GPU)
    integer(c_int32_t) function tp_field_deriv_noalloc(lxi, let, lze, nbc2, nn, src, dst, stream) &
        bind(C, name="tp_field_deriv_noalloc")
      import :: c_ptr, c_int32_t
      integer(c_int32_t), value :: lxi, let, lze, nn
      integer(c_int32_t), intent(in) :: nbc2(2)
      type(c_ptr), value :: src, dst, stream
    end function
    integer(c_int32_t) function tp_get_option(ctx, option, value) bind(C, name="tp_get_option")
      import :: c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: option
      integer(c_int64_t), intent(out) :: value
    end function
    integer(c_int32_t) function tp_server_set_staging(srv, ahead) bind(C, name="tp_server_set_staging")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: srv
      integer(c_int32_t), value :: ahead
    end function
    integer(c_int32_t) function tp_server_stats2(srv, out6) bind(C, name="tp_server_stats2")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: srv
      real(c_double), intent(out) :: out6(6)
    end function
    ! move set-up cases from other contexts into ctx
    integer(c_int32_t) function tp_adopt_cases(ctx, n, srcs, case_ids, new_ids) bind(C, name="tp_adopt_cases")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
      type(c_ptr), intent(in) :: srcs(*)
      integer(c_int32_t), intent(in) :: case_ids(*)
      integer(c_int32_t), intent(out) :: new_ids(*)
    end function
    ! the reference's synthetic eddies (inflow 3)
    integer(c_int32_t) function tp_eddy_params(nits, randnum, deck, out10) bind(C, name="tp_eddy_params")
      import :: c_int32_t, c_double
      integer(c_int32_t), value :: nits
      real(c_double), intent(in) :: randnum(*), deck(18)
      real(c_double), intent(out) :: out10(*)
    end function
    integer(c_int32_t) function tp_set_inflow_eddies(ctx, case_id, nits, eddies10) &
        bind(C, name="tp_set_inflow_eddies")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, nits
      real(c_double), intent(in) :: eddies10(*)
    end function
    ! the checks of tp_add_case without a context
    integer(c_int32_t) function tp_check_case(nblocks, dims, nbc, ncd, njct, jct, jvel, iparams) &
        bind(C, name="tp_check_case")
      import :: c_int32_t
      integer(c_int32_t), value :: nblocks, njct
      integer(c_int32_t), intent(in) :: dims(*), nbc(*), ncd(*), jct(*), jvel(*), iparams(*)
    end function
    integer(c_int32_t) function tp_step(ctx, nsteps) bind(C, name="tp_step")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int32_t) function tp_step_graph(ctx, nsteps) bind(C, name="tp_step_graph")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int32_t) function tp_sync(ctx) bind(C, name="tp_sync")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function tp_set_option(ctx, option, value) bind(C, name="tp_set_option")
      import :: c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: option
      integer(c_int64_t), value :: value
    end function
    integer(c_int32_t) function tp_autotune(ctx, reps, chosen, us) bind(C, name="tp_autotune")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: reps
      integer(c_int32_t), intent(out) :: chosen(6)
      real(c_double), intent(out) :: us(6)
    end function
    integer(c_int32_t) function tp_col_config(ctx, chunks6) bind(C, name="tp_col_config")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: chunks6(6)
    end function
    integer(c_int32_t) function tp_profile(ctx, on) bind(C, name="tp_profile")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int32_t) function tp_profile_classes(ctx, mask) bind(C, name="tp_profile_classes")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mask
    end function
    integer(c_int32_t) function tp_profile_read(ctx, kclass, ms, launches, bytes, reset) &
        bind(C, name="tp_profile_read")
      import :: c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kclass, reset
      real(c_double), intent(out) :: ms, bytes
      integer(c_int64_t), intent(out) :: launches
    end function
    integer(c_int32_t) function tp_info(ctx, out4) bind(C, name="tp_info")
      import :: c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: out4(4)
    end function
    integer(c_int32_t) function tp_inflow_modes(nmodes, seed, tu, lturb, uoo3, slit, span, out8) &
        bind(C, name="tp_inflow_modes")
      import :: c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: nmodes
      integer(c_int64_t), value :: seed
      real(c_double), value :: tu, lturb, slit, span
      real(c_double), intent(in) :: uoo3(3)
      real(c_double), intent(out) :: out8(*)
    end function
    integer(c_int32_t) function tp_gust_mode(tu, kr, uoo3, out8, slit) bind(C, name="tp_gust_mode")
      import :: c_int32_t, c_double
      real(c_double), value :: tu, kr
      real(c_double), intent(in) :: uoo3(3)
      real(c_double), intent(out) :: out8(8), slit
    end function
    ! the reference's file formats
    integer(c_int32_t) function tp_set_parents(ctx, case_id, nparent, parent, start3, pdims3) &
        bind(C, name="tp_set_parents")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, nparent
      integer(c_int32_t), intent(in) :: parent(*), start3(*), pdims3(*)
    end function
    integer(c_int32_t) function tp_get_parents(ctx, case_id, nparent, parent, start3, pdims3) &
        bind(C, name="tp_get_parents")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      integer(c_int32_t), intent(out) :: nparent, parent(*), start3(*), pdims3(*)
    end function
    integer(c_int32_t) function tp_get_grid(ctx, case_id, block, xyz, count) bind(C, name="tp_get_grid")
      import :: c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, block
      real(c_double), intent(out) :: xyz(*)
      integer(c_int64_t), value :: count
    end function
    integer(c_int32_t) function tp_case_dims(ctx, case_id, dims3) bind(C, name="tp_case_dims")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      integer(c_int32_t), intent(out) :: dims3(*)
    end function
    integer(c_int32_t) function tp_case_params(ctx, case_id, iparams, dparams) bind(C, name="tp_case_params")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      integer(c_int32_t), intent(out) :: iparams(16)
      real(c_double), intent(out) :: dparams(32)
    end function
    integer(c_int32_t) function tp_write_restart(ctx, case_id, rundir) bind(C, name="tp_write_restart")
      import :: c_ptr, c_int32_t, c_char
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      character(kind=c_char), intent(in) :: rundir(*)
    end function
    integer(c_int32_t) function tp_read_restart(ctx, case_id, rundir) bind(C, name="tp_read_restart")
      import :: c_ptr, c_int32_t, c_char
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      character(kind=c_char), intent(in) :: rundir(*)
    end function
    integer(c_int32_t) function tp_write_grid(ctx, case_id, rundir) bind(C, name="tp_write_grid")
      import :: c_ptr, c_int32_t, c_char
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      character(kind=c_char), intent(in) :: rundir(*)
    end function
    integer(c_int32_t) function tp_read_grid(ctx, case_id, rundir) bind(C, name="tp_read_grid")
      import :: c_ptr, c_int32_t, c_char
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      character(kind=c_char), intent(in) :: rundir(*)
    end function
    integer(c_int32_t) function tp_write_plot3d(ctx, case_id, rundir, what) bind(C, name="tp_write_plot3d")
      import :: c_ptr, c_int32_t, c_char
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, what
      character(kind=c_char), intent(in) :: rundir(*)
    end function
    ! surface statistics
    integer(c_int32_t) function tp_stats_start(ctx, case_id, every) bind(C, name="tp_stats_start")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, every
    end function
    integer(c_int32_t) function tp_stats_forces(ctx, case_id, out8) bind(C, name="tp_stats_forces")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      real(c_double), intent(out) :: out8(8)
    end function
    integer(c_int64_t) function tp_stats_wall_count(ctx, case_id, block) bind(C, name="tp_stats_wall_count")
      import :: c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, block
    end function
    integer(c_int32_t) function tp_stats_wall(ctx, case_id, block, mean_p, rms_p, area, wnor3, count) &
        bind(C, name="tp_stats_wall")
      import :: c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id, block
      real(c_double), intent(out) :: mean_p(*), rms_p(*), area(*), wnor3(*)
      integer(c_int64_t), value :: count
    end function
    ! cases joining and leaving a running context
    integer(c_int32_t) function tp_case_bytes(ctx, case_id, bytes) bind(C, name="tp_case_bytes")
      import :: c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      integer(c_int64_t), intent(out) :: bytes
    end function
    integer(c_int32_t) function tp_admit(ctx, case_id) bind(C, name="tp_admit")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
    end function
    integer(c_int32_t) function tp_admit_cases(ctx, n, case_ids, status) bind(C, name="tp_admit_cases")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: case_ids(*)
      integer(c_int32_t), intent(out) :: status(*)
    end function
    integer(c_int32_t) function tp_remove_cases(ctx, n, case_ids) bind(C, name="tp_remove_cases")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: case_ids(*)
    end function
    integer(c_int32_t) function tp_remove_case(ctx, case_id) bind(C, name="tp_remove_case")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
    end function
    integer(c_int32_t) function tp_case_info(ctx, case_id, out4) bind(C, name="tp_case_info")
      import :: c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: case_id
      integer(c_int64_t), intent(out) :: out4(4)
    end function
    ! request loop
    integer(c_int32_t) function tp_server_create(budget_bytes, max_active, use_graph, srv) &
        bind(C, name="tp_server_create")
      import :: c_ptr, c_int32_t, c_int64_t
      integer(c_int64_t), value :: budget_bytes
      integer(c_int32_t), value :: max_active, use_graph
      type(c_ptr), intent(out) :: srv
    end function
    integer(c_int32_t) function tp_server_destroy(srv) bind(C, name="tp_server_destroy")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: srv
    end function
    type(c_ptr) function tp_server_context(srv) bind(C, name="tp_server_context")
      import :: c_ptr
      type(c_ptr), value :: srv
    end function
    integer(c_int32_t) function tp_server_set_fill(srv, points) bind(C, name="tp_server_set_fill")
      import :: c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: srv
      integer(c_int64_t), value :: points
    end function
    type(c_ptr) function tp_server_error(srv) bind(C, name="tp_server_error")
      import :: c_ptr
      type(c_ptr), value :: srv
    end function
    real(c_double) function tp_server_clock(srv) bind(C, name="tp_server_clock")
      import :: c_ptr, c_double
      type(c_ptr), value :: srv
    end function
    integer(c_int32_t) function tp_server_submit(srv, nblocks, dims, nbc, ncd, nsz, njct, jct, jvel, &
                                                 iparams, dparams, xyz, nsteps, t_arrival, req) &
        bind(C, name="tp_server_submit")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: srv
      integer(c_int32_t), value :: nblocks, njct, nsteps
      integer(c_int32_t), intent(in) :: dims(*), nbc(*), ncd(*), nsz(*), jct(*), jvel(*), iparams(16)
      real(c_double), intent(in) :: dparams(32)
      type(c_ptr), intent(in) :: xyz(*)
      real(c_double), value :: t_arrival
      integer(c_int32_t), intent(out) :: req
    end function
    integer(c_int32_t) function tp_server_iterate(srv, max_steps, status6) bind(C, name="tp_server_iterate")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: srv
      integer(c_int32_t), value :: max_steps
      integer(c_int32_t), intent(out) :: status6(6)
    end function
    integer(c_int32_t) function tp_server_run(srv, max_steps) bind(C, name="tp_server_run")
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: srv
      integer(c_int32_t), value :: max_steps
    end function
    integer(c_int32_t) function tp_server_result(srv, req, qa, count, info6) bind(C, name="tp_server_result")
      import :: c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: srv
      integer(c_int32_t), value :: req
      real(c_double), intent(out) :: qa(*)
      integer(c_int64_t), value :: count
      real(c_double), intent(out) :: info6(6)
    end function
    integer(c_int64_t) function tp_server_result_count(srv, req) bind(C, name="tp_server_result_count")
      import :: c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: srv
      integer(c_int32_t), value :: req
    end function
    integer(c_int32_t) function tp_server_stats(srv, out16) bind(C, name="tp_server_stats")
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: srv
      real(c_double), intent(out) :: out16(16)
    end function
    pure integer(c_int64_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_int64_t
      type(c_ptr), value, intent(in) :: s
    end function
  end interface

contains

  ! message of the last failed call on this thread
  function tp_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = tp_last_error()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

  ! stop with the library's message unless rc is TP_OK
  subroutine tp_check(rc, where)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in), optional :: where
    if (rc == TP_OK) return
    if (present(where)) then
      write(*, '(a,": ",a)') where, tp_error_message()
    else
      write(*, '(a)') tp_error_message()
    end if
    error stop 1
  end subroutine

end module tandem_hip
