! tam_bind.f90 -- ISO_C_BINDING interface to libtam.so (include/tam.h).
!
! The derived types are interoperable with the C structs (bind(C) gives them
! the C layout, padding included); abi_check.f90 compares their c_sizeof
! with tam_struct_size.  Every function returns 0 or a negative TAM_ERR_*
! code; tam_check turns a failure into `error stop` with the library's
! message, as the reference stops on errors (3dFD.f90:179-182).
module tam_bind

    use iso_c_binding
    implicit none

    integer(c_int), parameter :: TAM_OK = 0, TAM_ERR_ARG = -1, TAM_ERR_HIP = -2, &
                                 TAM_ERR_NEGATIVE_TEMP = -3, TAM_ERR_STATE = -4, &
                                 TAM_ERR_NOMEM = -5, TAM_ERR_NODEVICE = -6
    integer(c_int32_t), parameter :: TAM_BEAM_TOPHAT = 0, TAM_BEAM_GAUSSIAN = 1
    integer(c_int32_t), parameter :: TAM_THERMAL_WATER = 0, TAM_THERMAL_TABLE = 1
    integer(c_int32_t), parameter :: TAM_FIELD_TEMP = 0, TAM_FIELD_JMEAN = 1, TAM_FIELD_TALLY = 2, &
                                     TAM_FIELD_RHOKAP = 3, TAM_FIELD_TISSUE = 4, TAM_FIELD_THRESTIME = 5, &
                                     TAM_FIELD_WATER = 6, TAM_FIELD_Q = 7, TAM_FIELD_KAPPA = 8, &
                                     TAM_FIELD_DENSITY = 9, TAM_FIELD_HEATCAP = 10, TAM_FIELD_COEFF = 11, &
                                     TAM_FIELD_MAT = 12

    type, bind(C) :: tam_material
        real(c_double) :: mua, mus, g, n
    end type tam_material

    type, bind(C) :: tam_params
        integer(c_int32_t) :: nx, ny, nz
        real(c_double)     :: xmax, ymax, zmax
        real(c_double)     :: dx_m, dy_m, dz_m, delt, dscale
        real(c_double)     :: volume_voxel, mass_voxel, qvapor, t_ref
        integer(c_int32_t) :: boil, robin_top
        real(c_double)     :: h_top, t_amb
        integer(c_int32_t) :: thermal_mode
        real(c_double)     :: mu_water, mu_protein, ablate_K
        real(c_double)     :: arr_A, arr_dE, arr_R, arr_Tlo, arr_Thi, arr_thr(3)
        integer(c_int32_t) :: damage_every_step, beam
        real(c_double)     :: beam_radius, n_ext_zlo, n_ext_zhi
        integer(c_int32_t) :: device, heat_fp64, removed_inert
    end type tam_params

    type, bind(C) :: tam_mc_stats
        real(c_double) :: packets, absorbed, specular, escaped(6), missed
        real(c_double) :: crossings, interactions, kernel_ms, killed, tally_adds, box_adds
    end type tam_mc_stats

    type, bind(C) :: tam_pulse
        real(c_double)     :: time, pulseCount, repetitionCount, repetitionRate_1, realPulseLength, laserOn
        integer(c_int32_t) :: laser_flag, pulsesDone, pulsesToDo, use_pulse
    end type tam_pulse

    type, bind(C) :: tam_timing
        real(c_double) :: mc_ms, normalise_ms, heat_ms, arrhenius_ms, rebuild_ms
        real(c_double) :: heat_steps, heat_step_ms, class_mode, tile, kz, cycle_ms
    end type tam_timing

    interface
        integer(c_int) function tam_abi_version() bind(C, name="tam_abi_version")
            import :: c_int
        end function

        integer(c_int64_t) function tam_struct_size(which) bind(C, name="tam_struct_size")
            import :: c_int64_t, c_int32_t
            integer(c_int32_t), value :: which
        end function

        integer(c_int) function tam_create(p, ctx) bind(C, name="tam_create")
            import :: c_int, c_ptr, tam_params
            type(tam_params), intent(in) :: p
            type(c_ptr), intent(out)     :: ctx
        end function

        subroutine tam_destroy(ctx) bind(C, name="tam_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine

        type(c_ptr) function tam_last_error(ctx) bind(C, name="tam_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function

        subroutine tam_pulse_advance(pulse, delt) bind(C, name="tam_pulse_advance")
            import :: tam_pulse, c_double
            type(tam_pulse), intent(inout) :: pulse
            real(c_double), value          :: delt
        end subroutine

        integer(c_int) function tam_set_optics(ctx, mat, table, nmat, rhokap) bind(C, name="tam_set_optics")
            import :: c_int, c_ptr, c_int8_t, c_int32_t, tam_material
            type(c_ptr), value                 :: ctx
            integer(c_int8_t), intent(in)      :: mat(*)
            type(tam_material), intent(in)     :: table(*)
            integer(c_int32_t), value          :: nmat
            type(c_ptr), value                 :: rhokap
        end function

        integer(c_int) function tam_set_thermal(ctx, temp, kappa, density, heatcap, q, w) &
                bind(C, name="tam_set_thermal")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value         :: ctx
            real(c_double), intent(in) :: temp(*), kappa(*), density(*), heatcap(*)
            type(c_ptr), value         :: q, w
        end function

        integer(c_int) function tam_mc_run(ctx, npackets, seed, stream, accumulate, stats) &
                bind(C, name="tam_mc_run")
            import :: c_int, c_ptr, c_int64_t, c_int32_t, tam_mc_stats
            type(c_ptr), value                :: ctx
            integer(c_int64_t), value         :: npackets, seed
            integer(c_int32_t), value         :: stream, accumulate
            type(tam_mc_stats), intent(inout) :: stats
        end function

        integer(c_int) function tam_normalise(ctx, scale) bind(C, name="tam_normalise")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: ctx
            real(c_double), value :: scale
        end function

        integer(c_int) function tam_heat_run(ctx, loops, pulse) bind(C, name="tam_heat_run")
            import :: c_int, c_ptr, c_int32_t, tam_pulse
            type(c_ptr), value             :: ctx
            integer(c_int32_t), value      :: loops
            type(tam_pulse), intent(inout) :: pulse
        end function

        integer(c_int) function tam_arrhenius(ctx, delt, time) bind(C, name="tam_arrhenius")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: ctx
            real(c_double), value :: delt, time
        end function

        integer(c_int) function tam_rebuild(ctx) bind(C, name="tam_rebuild")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function tam_cycle(ctx, npackets, seed, stream, scale, loops, arr_delt, pulse, stats) &
                bind(C, name="tam_cycle")
            import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double, tam_pulse
            type(c_ptr), value             :: ctx
            integer(c_int64_t), value      :: npackets, seed
            integer(c_int32_t), value      :: stream, loops
            real(c_double), value          :: scale, arr_delt
            type(tam_pulse), intent(inout) :: pulse
            type(c_ptr), value             :: stats
        end function

        integer(c_int) function tam_download(ctx, field, out) bind(C, name="tam_download")
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value          :: ctx
            integer(c_int32_t), value   :: field
            real(c_double), intent(out) :: out(*)
        end function

        integer(c_int) function tam_upload(ctx, field, in) bind(C, name="tam_upload")
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value         :: ctx
            integer(c_int32_t), value  :: field
            real(c_double), intent(in) :: in(*)
        end function

        integer(c_int) function tam_get_timing(ctx, t) bind(C, name="tam_get_timing")
            import :: c_int, c_ptr, tam_timing
            type(c_ptr), value              :: ctx
            type(tam_timing), intent(inout) :: t
        end function

        integer(c_int) function tam_synchronize(ctx) bind(C, name="tam_synchronize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        ! spectral helpers (ch_opt.f90:28-152): table is the reference's
        ! array(len, 2); errors through tam_message(c_null_ptr)
        integer(c_int) function tam_spectrum_cdf(table, len, cdf) bind(C, name="tam_spectrum_cdf")
            import :: c_int, c_int32_t, c_double
            real(c_double), intent(in)    :: table(*)
            integer(c_int32_t), value     :: len
            real(c_double), intent(inout) :: cdf(*)
        end function

        integer(c_int) function tam_spectrum_sample(device, table, cdf, len, n, seed, stream, wave) &
                bind(C, name="tam_spectrum_sample")
            import :: c_int, c_int32_t, c_int64_t, c_double
            integer(c_int32_t), value   :: device, len, stream
            real(c_double), intent(in)  :: table(*), cdf(*)
            integer(c_int64_t), value   :: n, seed
            real(c_double), intent(out) :: wave(*)
        end function

        integer(c_int) function tam_spectrum_interp(device, table, len, x, n, y) bind(C, name="tam_spectrum_interp")
            import :: c_int, c_int32_t, c_int64_t, c_double
            integer(c_int32_t), value   :: device, len
            real(c_double), intent(in)  :: table(*), x(*)
            integer(c_int64_t), value   :: n
            real(c_double), intent(out) :: y(*)
        end function

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

contains

    ! The library's last error message as a Fortran string.
    function tam_message(ctx) result(msg)
        type(c_ptr), intent(in)       :: ctx
        character(len=:), allocatable :: msg
        type(c_ptr)                   :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: i, n

        p = tam_last_error(ctx)
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        end do
    end function tam_message

    ! Stop with the library's message when rc reports a failure.
    subroutine tam_check(rc, ctx, what)
        integer(c_int), intent(in)   :: rc
        type(c_ptr), intent(in)      :: ctx
        character(len=*), intent(in) :: what

        if (rc /= TAM_OK) then
            write(*, '(a, a, i0, a, a)') what, ": tam error ", rc, ": ", tam_message(ctx)
            error stop 1
        end if
    end subroutine tam_check

end module tam_bind
