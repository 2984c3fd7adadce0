! tam_host.f90 -- host side of the reference run around the C-ABI:
!   * input.params, the 14 positional records of mcpolar.f90:79-94;
!   * the reference-default tissue and initThermalCoeff's scalars
!     (3dFD.f90:233-309, fits of thermalConst_mod.f90:11-87);
!   * the laser power functions (3dFD.f90:365-421) and the run-length rule
!     of mcpolar.f90:133-140;
!   * the output files of writer (writer.f90:6-81): raw FP64 column-major
!     streams named from the run parameters (utils.f90 str formatting).
module tam_host

    use iso_c_binding
    use tam_bind
    implicit none

    type :: run_input
        integer          :: nphotons, loops, pulsesToDo
        double precision :: xmax, ymax, zmax, n1, n2, total_time, repetitionRate_1, power
        double precision :: energyPerPixel, ablateTemp
        character(len=32):: pulsetype
    end type run_input

    double precision, parameter :: water0 = 0.75d0        ! initial water fraction (thermalConst_mod.f90:6)
    double precision, parameter :: latent = 2256d3        ! J/kg, latent heat of water
    double precision, parameter :: spots = 81.d0          ! spotsPerRow * spotsPerCol (constants.f90)

contains

    subroutine read_input(path, inp)
        character(len=*), intent(in) :: path
        type(run_input), intent(out) :: inp
        integer :: u, ios

        open(newunit=u, file=path, status="old", action="read", iostat=ios)
        if (ios /= 0) then
            write(*, '(a, a)') "cannot open ", trim(path)
            error stop 1
        end if
        ! list-directed reads: the trailing comments end at the record, and the
        ! "/" in "10000000./1." ends the value as in the reference's read
        read(u, *) inp%nphotons
        read(u, *) inp%xmax
        read(u, *) inp%ymax
        read(u, *) inp%zmax
        read(u, *) inp%n1
        read(u, *) inp%n2
        read(u, *) inp%total_time
        read(u, *) inp%loops
        read(u, *) inp%repetitionRate_1
        read(u, *) inp%power
        read(u, *) inp%energyPerPixel
        read(u, *) inp%ablateTemp
        read(u, *) inp%pulsesToDo
        read(u, *) inp%pulsetype
        close(u)
    end subroutine read_input

    ! skin property fits of the water fraction w (protein fraction 1 - w0)
    pure double precision function skin_density(w)
        double precision, intent(in) :: w
        skin_density = 1000.d0 / (w + 0.649d0 * (1.d0 - water0))
    end function
    pure double precision function skin_heatcap(w)
        double precision, intent(in) :: w
        skin_heatcap = 1000.d0 * (4.2d0 * w + 1.09d0 * (1.d0 - water0))
    end function
    pure double precision function skin_cond(w, rho)
        double precision, intent(in) :: w, rho
        skin_cond = rho * (6.28d-4 * w + 1.17d-4 * (1.d0 - water0))
    end function
    pure double precision function air_cond(t)
        double precision, intent(in) :: t
        air_cond = -0.188521d0 * exp(-0.000367259d0 * (t - 273.15d0)) + 0.212453d0
    end function

    ! The reference-default tissue on an n^3 grid: tam_params, the halo
    ! temperature and property arrays (0:n+1)^3, mu_t (rhokap) and the one
    ! material (mu_a = 680 /cm pure absorber, n2).  Returns the pulse length
    ! (energy per pixel over the power, 3dFD.f90:287).
    subroutine reference_setup(inp, n, p, temp, kappa, density, heatcap, rhokap, mat, table, pulse_length)
        type(run_input), intent(in)   :: inp
        integer, intent(in)           :: n
        type(tam_params), intent(out) :: p
        double precision, intent(out) :: temp(0:n+1, 0:n+1, 0:n+1), kappa(0:n+1, 0:n+1, 0:n+1)
        double precision, intent(out) :: density(0:n+1, 0:n+1, 0:n+1), heatcap(0:n+1, 0:n+1, 0:n+1)
        double precision, intent(out) :: rhokap(0:n+1, 0:n+1, 0:n+1)
        integer(c_int8_t), intent(out):: mat(n, n, n)
        type(tam_material), intent(out) :: table(1)
        double precision, intent(out) :: pulse_length
        double precision :: rho, hc, k, alpha, constd

        rho = skin_density(water0)
        hc = skin_heatcap(water0)
        k = skin_cond(water0, rho)
        alpha = k / (rho * hc)

        p = tam_params(nx=n, ny=n, nz=n, xmax=inp%xmax, ymax=inp%ymax, zmax=inp%zmax, &
                       dx_m=0.d0, dy_m=0.d0, dz_m=0.d0, delt=0.d0, dscale=0.5d0, &
                       volume_voxel=0.d0, mass_voxel=0.d0, qvapor=0.d0, t_ref=278.d0, &
                       boil=1, robin_top=0, h_top=0.d0, t_amb=0.d0, thermal_mode=TAM_THERMAL_WATER, &
                       mu_water=510.d0, mu_protein=170.d0, ablate_K=inp%ablateTemp + 273.d0, &
                       arr_A=3.1d98, arr_dE=6.3d5, arr_R=8.314d0, arr_Tlo=43.d0 + 273.d0, &
                       arr_Thi=100.d0 + 273.d0, arr_thr=[0.53d0, 1.d0, 1.d4], damage_every_step=0, &
                       beam=TAM_BEAM_TOPHAT, beam_radius=125d-4, n_ext_zlo=0.d0, n_ext_zhi=0.d0, &
                       device=0, heat_fp64=1, removed_inert=0)
        ! FD node spacing over n+2 nodes and the explicit limit of the
        ! half-factor scheme (3dFD.f90:249-281)
        p%dx_m = (2.d0 * inp%xmax * 1.d-2) / (n + 2.d0)
        p%dy_m = (2.d0 * inp%ymax * 1.d-2) / (n + 2.d0)
        p%dz_m = (2.d0 * inp%zmax * 1.d-2) / (n + 2.d0)
        constd = 1.d0 / p%dx_m**2 + 1.d0 / p%dy_m**2 + 1.d0 / p%dz_m**2
        p%delt = 1.d0 / (alpha * constd)
        ! MC voxel volume and mass, latent heat budget (3dFD.f90:289-291)
        p%volume_voxel = (2.d0 * inp%xmax * 1.d-2 / n) * (2.d0 * inp%ymax * 1.d-2 / n) * &
                         (2.d0 * inp%zmax * 1.d-2 / n)
        p%mass_voxel = rho * p%volume_voxel
        p%qvapor = latent * p%mass_voxel

        ! halo: air conductivity at 25 C, skin density and heat capacity;
        ! interior skin (3dFD.f90:262-278)
        kappa = air_cond(25.d0 + 273.d0)
        kappa(1:n, 1:n, 1:n) = k
        density = rho
        heatcap = hc
        ! 5 C inside and on the sides, 25 C below and above (mcpolar.f90:123-129)
        temp = 5.d0 + 273.d0
        temp(:, :, 0) = 25.d0 + 273.d0
        temp(:, :, n + 1) = 25.d0 + 273.d0
        rhokap = 0.d0
        rhokap(1:n, 1:n, 1:n) = 680.d0
        mat = 0_c_int8_t
        table(1) = tam_material(mua=680.d0, mus=0.d0, g=0.9d0, n=inp%n2)
        pulse_length = inp%energyPerPixel * 1.d-3 * spots / inp%power
    end subroutine reference_setup

    ! Laser power at the current time (3dFD.f90:365-421).  ramp_down carries
    ! the triangular pulse's state between calls.
    double precision function laser_power(inp, pulse, pulse_length, ramp_down) result(pw)
        type(run_input), intent(in)  :: inp
        type(tam_pulse), intent(in)  :: pulse
        double precision, intent(in) :: pulse_length
        logical, intent(inout)       :: ramp_down
        double precision :: fwhm, mu, sig, slope

        select case (trim(inp%pulsetype))
        case ("gaussian")
            fwhm = 2.d0 * sqrt(2.d0 * log(2.d0))
            mu = fwhm * pulse_length
            sig = pulse_length / fwhm
            pw = inp%power * exp(-(pulse%time - mu)**2 / (2.d0 * sig**2))
        case ("tophat")
            pw = merge(inp%power, 0.d0, pulse%laser_flag /= 0)
        case ("triangular")
            slope = inp%power / pulse_length
            if (pulse%laser_flag == 0) then
                pw = 0.d0
            else
                if (.not. ramp_down .and. pulse%time >= pulse_length) ramp_down = .true.
                if (ramp_down) then
                    pw = max(0.d0, -slope * pulse%time + 2.d0 * inp%power)
                else
                    pw = slope * pulse%time
                end if
            end if
        case default
            error stop "no pulse type"
        end select
    end function laser_power

    ! The full length of the laser pulse and the run's end time
    ! (3dFD.f90:293-306, mcpolar.f90:133-140).
    subroutine run_length(inp, pulse_length, delt, real_pulse, total_time)
        type(run_input), intent(in)   :: inp
        double precision, intent(in)  :: pulse_length, delt
        double precision, intent(out) :: real_pulse, total_time

        select case (trim(inp%pulsetype))
        case ("tophat")
            real_pulse = pulse_length
        case ("gaussian")
            real_pulse = 20000.d0 * pulse_length
        case ("triangular")
            real_pulse = 2.d0 * pulse_length
        case default
            error stop "no pulse type"
        end select
        total_time = inp%total_time
        if (trim(inp%pulsetype) == "gaussian") then
            total_time = 2.d0 * pulse_length * (2.d0 * sqrt(2.d0 * log(2.d0)))
            real_pulse = total_time
        else if (int(total_time / delt) <= int(real_pulse / delt)) then
            total_time = delt * (real_pulse / delt + 2000.d0)
        end if
    end subroutine run_length

    ! heat_sim_3D's guard (3dFD.f90:102-105): a pulse shorter than one
    ! explicit step shortens the step to a hundredth of the pulse.  The
    ! reference applies it on the first heat call, after mcpolar has set the
    ! run length from the original step (mcpolar.f90:133-140), and
    ! initThermalCoeff has already set coeff = delt/(rho c) from the original
    ! step; setupThermalCoeff rebuilds coeff from the new one after the first
    ! cycle (3dFD.f90:339-345).  Returns .true. when the step changed.
    logical function step_guard(pulse_length, delt)
        double precision, intent(in)    :: pulse_length
        double precision, intent(inout) :: delt

        step_guard = pulse_length < delt
        if (step_guard) then
            print *, "pulselength smaller than timestep, adjusting..."
            delt = pulse_length / 100.d0
        end if
    end function step_guard

    ! ---- file names: the reference formats numbers through f100.16 / I100.1
    ! records, left-adjusts them and keeps the leading `w` characters (with
    ! zero padding on the left for a short integer)
    function int_text(i, w) result(s)
        integer, intent(in)           :: i
        integer, intent(in), optional :: w
        character(len=:), allocatable :: s
        character(len=100) :: buf

        write(buf, '(i0)') i
        s = trim(adjustl(buf))
        if (present(w)) then
            if (len(s) <= w) then
                s = repeat("0", w - len(s)) // s
            else
                s = s(1:w)
            end if
        end if
    end function int_text

    function real_text(x, w) result(s)
        double precision, intent(in)  :: x
        integer, intent(in)           :: w
        character(len=:), allocatable :: s
        character(len=100) :: buf

        write(buf, '(f100.16)') x
        s = trim(adjustl(buf))
        s = trim(adjustl(s(1:min(w, len(s)))))
    end function real_text

    function run_tag(inp, n) result(tag)
        type(run_input), intent(in)   :: inp
        integer, intent(in)           :: n
        character(len=:), allocatable :: tag

        tag = int_text(int(inp%power)) // "w-" // int_text(n) // "-" // real_text(inp%ablateTemp, 3) // "-" // &
              int_text(int(inp%energyPerPixel), 3) // "-" // real_text(inp%xmax, 5) // "-" // &
              real_text(inp%ymax, 5) // "-" // real_text(inp%zmax, 5) // ".dat"
    end function run_tag

    subroutine write_stream(path, a, n)
        character(len=*), intent(in) :: path
        integer, intent(in)          :: n
        double precision, intent(in) :: a(n)
        integer :: u

        open(newunit=u, file=path, access="stream", form="unformatted", status="replace")
        write(u) a
        close(u)
    end subroutine write_stream

    ! writer.f90:6-81: jmean, rhokap (interior), temp - 273 (halo), water,
    ! tissue (-1 in the crater), and the three threshold-time fields.
    subroutine write_outputs(ctx, inp, n, outdir)
        type(c_ptr), intent(in)      :: ctx
        type(run_input), intent(in)  :: inp
        integer, intent(in)          :: n
        character(len=*), intent(in) :: outdir
        double precision, allocatable :: f(:), rk(:), halo(:), tis(:), thr(:)
        character(len=:), allocatable :: tag
        integer :: nv, nh, v

        nv = n * n * n
        nh = (n + 2)**3
        allocate(f(nv), rk(nh), halo(nh), tis(nv), thr(3 * nv))
        tag = run_tag(inp, n)

        call tam_check(tam_download(ctx, TAM_FIELD_JMEAN, f), ctx, "download jmean")
        call write_stream(outdir // "jmean-t" // tag, f, nv)

        call tam_check(tam_download(ctx, TAM_FIELD_RHOKAP, rk), ctx, "download rhokap")
        call interior(rk, f, n)
        call write_stream(outdir // "rhokap-t" // tag, f, nv)

        call tam_check(tam_download(ctx, TAM_FIELD_TEMP, halo), ctx, "download temp")
        call write_stream(outdir // "temp-t" // tag, halo - 273.d0, nh)

        call tam_check(tam_download(ctx, TAM_FIELD_WATER, f), ctx, "download water")
        call write_stream(outdir // "water-t" // tag, f, nv)

        ! no tissue is left to damage in the crater (mcpolar.f90:190-201)
        call tam_check(tam_download(ctx, TAM_FIELD_TISSUE, tis), ctx, "download tissue")
        call interior(rk, f, n)
        do v = 1, nv
            if (f(v) <= 0.1d0) tis(v) = -1.d0
        end do
        call write_stream(outdir // "tissue-t" // tag, tis, nv)

        call tam_check(tam_download(ctx, TAM_FIELD_THRESTIME, thr), ctx, "download threshold times")
        call write_stream(outdir // "/time-t-1-" // tag, thr(1:nv), nv)
        call write_stream(outdir // "/time-t-2-" // tag, thr(nv + 1:2 * nv), nv)
        call write_stream(outdir // "/time-t-3-" // tag, thr(2 * nv + 1:3 * nv), nv)
    end subroutine write_outputs

    ! interior (1:n)^3 of a halo array (0:n+1)^3, both column-major
    subroutine interior(h, out, n)
        integer, intent(in)           :: n
        double precision, intent(in)  :: h(0:n+1, 0:n+1, 0:n+1)
        double precision, intent(out) :: out(n, n, n)
        out = h(1:n, 1:n, 1:n)
    end subroutine interior

end module tam_host
