!!**************************************************************************************************
!! recurgreen_gpu.f90 -- namelist-compatible `tdistrib` driver whose sweep runs on the GPU.
!!
!! Usage: recurgreen_gpu settings.nml [more.nml ...]
!! Reads the reference's groups unchanged: &settings task (src/main.f90:45), &waveguide and
!! &tdistrib (src/main.f90:83-84), and the optional group
!!   &gpu  device, precision ('fp64' | 'mixed_fp32' | 'mixed_bf16x3'), batch, ny, dos_model, bintype /
!! that has no counterpart in the reference (a file without it parses as before).  The OpenMP loop of
!! compute_eigvals_omp (src/main.f90:134-179) becomes one batched device pass; the histogram chain
!! follows prepare_distrib (src/main.f90:223-251) and the CSV layout of save_distrib / save_raw_eigvals
!! (src/main.f90:257-328), without the TikZ step.
!!**************************************************************************************************
program recurgreen_gpu
    use, intrinsic :: iso_fortran_env, only: real64, error_unit, output_unit
    use recurgreen_hip
    implicit none
    character(len=300) :: arg
    integer :: i, funit, rstat
    logical :: fexist

    write (output_unit, '(a)') "====== This is RecurGreen-HIP (gfx950 sweep behind the RecurGreen v1.0 interface) ======"
    if (command_argument_count() == 0) write (error_unit, '(a)') "[ERROR] No settings file provided, doing nothing..."
    do i = 1, command_argument_count()
        call get_command_argument(i, arg)
        inquire(file=arg, exist=fexist)
        if (.not. fexist) then
            write (error_unit, '(a)') "[ERROR] File not found: " // trim(arg)
            cycle
        end if
        open(newunit=funit, file=arg, action='read', iostat=rstat)
        if (rstat /= 0) then
            write (error_unit, '(a)') "[ERROR] Cannot open file: " // trim(arg)
            cycle
        end if
        write (output_unit, '(a)') "[INFO] Reading file: " // trim(arg)
        call execute_file(funit)
        close(funit)
    end do

contains

    subroutine rewind_or_stop(funit, istat, optional_group)
        integer, intent(in) :: funit, istat
        logical, intent(in) :: optional_group
        if (istat == 0 .or. (optional_group .and. istat < 0)) then
            rewind(funit)
        else
            write (error_unit, '(a)') "[ERROR] Invalid entry in namelist, aborting..."
            stop errmsg_invalid_arg
        end if
    end subroutine

    subroutine execute_file(funit)
        integer, intent(in) :: funit
        character(len=50) :: task
        integer :: istat
        namelist /settings/ task
        read(unit=funit, nml=settings, iostat=istat)
        call rewind_or_stop(funit, istat, .false.)
        select case (task)
            case ("tdistrib")
                call task_tdistrib(funit)
            case ("wavefunction", "tprofile", "tprofile_group", "tpattern")
                write (error_unit, '(a)') "[ERROR] Task '" // trim(task) // "' is not part of the GPU sweep, aborting..."
                stop errmsg_invalid_arg
            case default
                write (error_unit, '(a)') "[ERROR] Unknown task '" // trim(task) // "', aborting..."
                stop errmsg_invalid_arg
        end select
    end subroutine

    subroutine task_tdistrib(funit)
        integer, intent(in) :: funit
        type(waveguide_gpu_t) :: wguide
        character(len=50) :: boundtype, precision, bintype
        real(wp) :: wol, aratio, exfac, dscat, dabso, napera, naperb, time, speed, missing
        integer :: nxdiso, nxfree, nseed, nbins, nthreads, istat, device, batch, ny, dos_model, prec, k
        logical :: save_eigvals
        integer(8) :: c0, c1, crate
        real(wp), allocatable :: tallvals(:), bins(:), histo(:), rho(:, :)
        namelist /waveguide/ boundtype, wol, exfac, dscat, dabso, aratio, nxdiso, nxfree, napera, naperb
        namelist /tdistrib/ nseed, nbins, nthreads, save_eigvals
        namelist /gpu/ device, precision, batch, ny, dos_model, bintype

        boundtype = "dirichlet"; wol = 4.3_wp; exfac = 1.0_wp; aratio = 1.0_wp; dscat = 1.0_wp; dabso = 0.0_wp
        nxdiso = 100; nxfree = 0; napera = 1.0_wp; naperb = 1.0_wp
        nseed = 1; nbins = 128; nthreads = 1; save_eigvals = .false.
        device = 0; precision = "fp64"; batch = 1024; ny = 0; dos_model = 0; bintype = "chebyshev"
        read(unit=funit, nml=waveguide, iostat=istat)
        call rewind_or_stop(funit, istat, .false.)
        read(unit=funit, nml=tdistrib, iostat=istat)
        call rewind_or_stop(funit, istat, .false.)
        read(unit=funit, nml=gpu, iostat=istat)
        call rewind_or_stop(funit, istat, .true.)
        select case (precision)
            case ("fp64"); prec = RG_FP64
            case ("mixed_fp32"); prec = RG_MIXED_FP32
            case ("mixed_bf16x3"); prec = RG_MIXED_BF16X3
            case default
                write (error_unit, '(a)') "[ERROR] Unknown precision '" // trim(precision) // "', aborting..."
                stop errmsg_invalid_arg
        end select
        if (nseed < 1 .or. nbins < 1) then
            write (error_unit, '(a)') "[ERROR] Invalid nseed or nbins, aborting..."
            stop errmsg_invalid_arg
        end if

        call wguide%set_options(ny, dos_model, prec)
        call wguide%init(boundtype, nxdiso, nxfree, wol, exfac, aratio, dscat, dabso, napera, naperb, &
                         batch_max=min(batch, nseed), device=device)
        call wguide%print_parameters(output_unit, "[INFO]")
        write (output_unit, '(3(a,i0),2a)') "[INFO] Nseed=", nseed, ", Nbins=", nbins, ", Device=", device, &
            ", Precision=", trim(precision)

        allocate(tallvals(nseed*wguide%get_neigv()))
        tallvals = 0.0_wp
        call system_clock(c0, crate)
        call wguide%compute_tdistrib_batch(1, nseed, tallvals)
        call system_clock(c1)
        time = real(c1 - c0, wp)/real(crate, wp)
        speed = nseed/time
        write (output_unit, '(a,f0.3,a,g0.6,a)') "[INFO] Done, computation time is ", time, " s (", speed, " seed/s)."

        !! prepare_distrib (src/main.f90:223-251)
        allocate(bins(nbins+1), histo(nbins), rho(nbins, 2))
        do k = 1, nbins+1
            if (bintype == "linear") then
                bins(k) = real(k-1, wp)/nbins
            else
                bins(k) = (1.0_wp - cos(acos(-1.0_wp)*(k-1)/nbins))/2
            end if
        end do
        call wguide%histogram(tallvals, bins, histo, missing)
        if (missing /= 0.0_wp) write (output_unit, '(a,f0.1,a)') "[WARN] There are ", missing, &
            " missing eigenvalues in the histogram, continuing anyway..."
        histo = histo + epsilon(histo(1))/10.
        histo = histo/(sum(histo)*(bins(2:nbins+1) - bins(1:nbins)))
        rho(:, 1) = (bins(1:nbins) + bins(2:nbins+1))/2
        rho(:, 2) = histo
        call save_files(wguide, rho, tallvals, nseed, nbins, time, speed, save_eigvals)
        call wguide%del()
    end subroutine

    subroutine save_files(wguide, rho, tallvals, nseed, nbins, time, speed, save_eigvals)
        type(waveguide_gpu_t), intent(in) :: wguide
        real(wp), intent(in) :: rho(:, :), tallvals(:), time, speed
        integer, intent(in) :: nseed, nbins
        logical, intent(in) :: save_eigvals
        character(len=200) :: filename
        integer :: outfp, k, ifile
        logical :: fexist
        call execute_command_line("mkdir -p out_gpu/distrib")
        do ifile = 1, 1000000   !! unique_filename (src/base_utils.f90:825-840)
            write (filename, '(a,i0,a)') "out_gpu/distrib/result_", ifile, ".distrib.csv"
            inquire(file=filename, exist=fexist)
            if (.not. fexist) exit
        end do
        open(newunit=outfp, file=filename, action='write')
        call wguide%print_parameters(outfp, "%%")
        write (outfp, '(2(a,i0),2(a,g0.6),a)') "%% Nseed=", nseed, ", Nbins=", nbins, ", Time=", time, &
            " s, Speed=", speed, " seed/s"
        write (outfp, '(a)') "tm, rho"
        do k = 1, size(rho, 1)
            write (outfp, '(g0.16,a,g0.16)') rho(k, 1), ", ", rho(k, 2)
        end do
        close(outfp)
        write (output_unit, '(a)') "[INFO] Result written in file: " // trim(filename)
        if (save_eigvals) then
            write (filename, '(a,i0,a)') "out_gpu/distrib/result_", ifile, ".eigvals.csv"
            open(newunit=outfp, file=filename, action='write')
            call wguide%print_parameters(outfp, "%%")
            write (outfp, '(a,i0,a)') "%% Transmission eigenvalues (", size(tallvals), " values)"
            do k = 1, size(tallvals)
                write (outfp, '(g0.16)') tallvals(k)
            end do
            close(outfp)
            write (output_unit, '(a)') "[INFO] Eigenvalues saved in file: " // trim(filename)
        end if
    end subroutine

end program
