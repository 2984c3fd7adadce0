! mcpolar_gpu.f90 -- the reference's driver (mcpolar.f90) on one MI355X.
!
! Same inputs (res/input.params), same time loop and same output files as
! mcpolar: while time <= total_time, a transport run of nphotons packets
! when the laser is on, normalised by the pulse power; `loops` heat steps;
! Arrhenius; setupThermalCoeff (mcpolar.f90:148-186); then writer.  Each
! loop iteration is one tam_cycle call (one HIP graph replay), and the host
! does not wait for the GPU between iterations.
!
! usage: mcpolar_gpu [resdir [outdir [n [max_cycles [inert]]]]]
!   resdir      directory holding input.params (default res/)
!   outdir      directory for the output files (default data/)
!   n           grid points per axis (default 80, constants.f90:12)
!   max_cycles  stop after this many cycles (default -1: run to total_time)
!   inert       1: no conduction through ablated voxels (default 0: they get
!               air properties as in 3dFD.f90:350-357, whose diffusivity is
!               unstable at the tissue's explicit time step once two ablated
!               voxels touch, and the run stops on T < 0 like the reference)
program mcpolar_gpu

    use iso_c_binding
    use tam_bind
    use tam_host
    implicit none

    type(run_input)    :: inp
    type(tam_params)   :: p
    type(tam_pulse)    :: pulse
    type(tam_material) :: table(1)
    type(tam_timing)   :: tm
    type(tam_mc_stats), target :: stats
    type(c_ptr)        :: ctx
    double precision, allocatable :: temp(:,:,:), kappa(:,:,:), density(:,:,:), heatcap(:,:,:)
    double precision, allocatable, target :: rhokap(:,:,:)
    double precision, allocatable :: coeff(:,:,:)
    integer(c_int8_t), allocatable :: mat(:,:,:)
    character(len=512) :: arg, resdir, outdir
    integer            :: n, max_cycles, cycles, nargs, inert
    integer(c_int)     :: rc
    integer(c_int64_t) :: npk, seed
    double precision   :: pulse_length, real_pulse, total_time, pw, scale, delt0
    logical            :: short_pulse
    logical            :: ramp_down
    integer(c_int64_t) :: t0, t1, rate

    resdir = "res/"
    outdir = "data/"
    n = 80
    max_cycles = -1
    inert = 0
    nargs = command_argument_count()
    if (nargs >= 1) call get_command_argument(1, resdir)
    if (nargs >= 2) call get_command_argument(2, outdir)
    if (nargs >= 3) then
        call get_command_argument(3, arg)
        read(arg, *) n
    end if
    if (nargs >= 4) then
        call get_command_argument(4, arg)
        read(arg, *) max_cycles
    end if
    if (nargs >= 5) then
        call get_command_argument(5, arg)
        read(arg, *) inert
    end if
    resdir = with_slash(resdir)
    outdir = with_slash(outdir)

    call read_input(trim(resdir) // "input.params", inp)
    allocate(temp(0:n+1, 0:n+1, 0:n+1), kappa(0:n+1, 0:n+1, 0:n+1), density(0:n+1, 0:n+1, 0:n+1), &
             heatcap(0:n+1, 0:n+1, 0:n+1), rhokap(0:n+1, 0:n+1, 0:n+1), mat(n, n, n))
    call reference_setup(inp, n, p, temp, kappa, density, heatcap, rhokap, mat, table, pulse_length)
    call run_length(inp, pulse_length, p%delt, real_pulse, total_time)
    delt0 = p%delt
    short_pulse = step_guard(pulse_length, p%delt)
    p%removed_inert = int(inert, c_int32_t)
    write(*, '(a, i0, a, i0, a)') "# photons per cycle ", inp%nphotons, ", grid ", n, "^3"
    write(*, '(a, es23.16, a, es23.16, a, es23.16)') "# delt ", p%delt, " total_time ", total_time, &
        " pulse ", real_pulse
    write(*, '(a, a)') "# outputs ", "jmean-t" // run_tag(inp, n)

    call tam_check(tam_create(p, ctx), c_null_ptr, "tam_create")
    call tam_check(tam_set_optics(ctx, mat, table, 1_c_int32_t, c_loc(rhokap)), ctx, "tam_set_optics")
    call tam_check(tam_set_thermal(ctx, temp, kappa, density, heatcap, c_null_ptr, c_null_ptr), ctx, &
                   "tam_set_thermal")
    if (short_pulse) then
        ! the first cycle's source coefficient keeps the original step
        allocate(coeff(0:n+1, 0:n+1, 0:n+1))
        coeff = 0.d0
        coeff(1:n, 1:n, 1:n) = delt0 / (density(1:n, 1:n, 1:n) * heatcap(1:n, 1:n, 1:n))
        call tam_check(tam_upload(ctx, TAM_FIELD_COEFF, coeff), ctx, "tam_upload")
        deallocate(coeff)
    end if

    pulse = tam_pulse(time=0.d0, pulseCount=0.d0, repetitionCount=0.d0, repetitionRate_1=inp%repetitionRate_1, &
                      realPulseLength=real_pulse, laserOn=1.d0, laser_flag=1, pulsesDone=0, &
                      pulsesToDo=inp%pulsesToDo, use_pulse=1)
    ramp_down = .false.
    ! the reference's ran2 seed (mcpolar.f90:97-98) keys the Philox streams
    seed = int(abs(-95648324), c_int64_t)

    call system_clock(t0, rate)
    cycles = 0
    do while (pulse%time <= total_time)
        if (max_cycles >= 0 .and. cycles >= max_cycles) exit
        npk = 0
        scale = 0.d0
        if (pulse%laser_flag /= 0) then
            npk = int(inp%nphotons, c_int64_t)
            pw = laser_power(inp, pulse, pulse_length, ramp_down)
            scale = (pw / spots) / (dble(inp%nphotons) * p%volume_voxel)   ! mcpolar.f90:174
        end if
        ! every 256th cycle waits for the GPU and checks for T < 0, where the
        ! reference stops (3dFD.f90:179-182)
        if (mod(cycles + 1, 256) == 0) then
            rc = tam_cycle(ctx, npk, seed, int(cycles, c_int32_t), scale, int(inp%loops, c_int32_t), &
                           p%delt, pulse, c_loc(stats))
        else
            rc = tam_cycle(ctx, npk, seed, int(cycles, c_int32_t), scale, int(inp%loops, c_int32_t), &
                           p%delt, pulse, c_null_ptr)
        end if
        if (rc /= TAM_OK) write(*, '(a, i0, a, es12.5, a)') "stopped in the 256 cycles up to ", cycles + 1, &
            " (t = ", pulse%time, " s)"
        call tam_check(rc, ctx, "tam_cycle")
        cycles = cycles + 1
    end do
    call tam_check(tam_synchronize(ctx), ctx, "tam_synchronize")
    call system_clock(t1)
    call tam_check(tam_get_timing(ctx, tm), ctx, "tam_get_timing")
    write(*, '(a, i0, a, f10.3, a, f10.6, a)') "cycles ", cycles, ", ", dble(t1 - t0) / dble(rate), &
        " s, simulated time ", pulse%time, " s"

    call write_outputs(ctx, inp, n, trim(outdir))
    call tam_destroy(ctx)

contains

    function with_slash(d) result(s)
        character(len=*), intent(in) :: d
        character(len=len(d)) :: s
        integer :: l

        s = d
        l = len_trim(d)
        if (l > 0 .and. l < len(d)) then
            if (d(l:l) /= "/") s(l + 1:l + 1) = "/"
        end if
    end function with_slash

end program mcpolar_gpu
