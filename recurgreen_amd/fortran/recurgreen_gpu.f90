!!**************************************************************************************************
!! recurgreen_gpu.f90 -- namelist-compatible task driver (tdistrib, tpattern, tprofile, tprofile_group,
!! wavefunction) whose sweeps run on the GPU.
!!
!! Usage: recurgreen_gpu settings.nml [more.nml ...]
!! Reads the reference's groups unchanged: &settings task (src/main.f90:45), &waveguide and
!! &tdistrib (src/main.f90:83-84), and the optional group
!!   &gpu  device, precision ('fp64' | 'mixed_fp32' | 'mixed_bf16x3'), batch, ny, dos_model, bintype,
!!         refine (tdistrib: FP64 refinement passes of G(L;0) for the mixed precisions, default 0) /
!! that has no counterpart in the reference (a file without it parses as before).  The OpenMP loop of
!! compute_eigvals_omp (src/main.f90:134-179) becomes one batched device pass; the histogram chain
!! follows prepare_distrib (src/main.f90:223-251) and the CSV layout of save_distrib / save_raw_eigvals
!! (src/main.f90:257-328), without the TikZ step.  Task `tpattern` (src/main.f90:735-901) reads
!! &tpattern nseed, imode, nthreads, probatype and averages |t|^2, |r|^2 per channel over batched
!! complex128 reflection/transmission passes.
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
            case ("tpattern")
                call task_tpattern(funit)
            case ("wavefunction")
                call task_wavefunction(funit)
            case ("tprofile")
                call task_tprofile(funit)
            case ("tprofile_group")
                call task_tprofile_group(funit)
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
        integer :: nxdiso, nxfree, nseed, nbins, nthreads, istat, device, batch, ny, dos_model, prec, k, refine
        logical :: save_eigvals
        integer(8) :: c0, c1, crate
        real(wp), allocatable :: tallvals(:), bins(:), histo(:), rho(:, :)
        namelist /waveguide/ boundtype, wol, exfac, dscat, dabso, aratio, nxdiso, nxfree, napera, naperb
        namelist /tdistrib/ nseed, nbins, nthreads, save_eigvals
        namelist /gpu/ device, precision, batch, ny, dos_model, bintype, refine

        boundtype = "dirichlet"; wol = 4.3_wp; exfac = 1.0_wp; aratio = 1.0_wp; dscat = 1.0_wp; dabso = 0.0_wp
        nxdiso = 100; nxfree = 0; napera = 1.0_wp; naperb = 1.0_wp
        nseed = 1; nbins = 128; nthreads = 1; save_eigvals = .false.
        device = 0; precision = "fp64"; batch = 1024; ny = 0; dos_model = 0; bintype = "chebyshev"; refine = 0
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
        if (refine > 0) call wguide%set_plan_option("refine", refine)
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
        call save_files(wguide, rho, tallvals, nseed, nbins, nthreads, time, speed, save_eigvals)
        call wguide%del()
    end subroutine

    !! task_tpattern (src/main.f90:735-856): the seed loop of compute_refl_trans_v1 (probatype
    !! 'incident', 'total') or compute_deposition (probatype 'deposition', from the first slice to the
    !! middle one, :764-765) runs as batched complex128 device passes of `batch` realizations; the
    !! per-channel probabilities are summed pass by pass and averaged (:838-840).  The &gpu group is
    !! read as in task_tdistrib; its precision and bintype do not apply (these matrices are complex128).
    subroutine task_tpattern(funit)
        integer, intent(in) :: funit
        type(waveguide_gpu_t) :: wguide
        character(len=50) :: boundtype, probatype, precision, bintype
        real(wp) :: wol, aratio, exfac, dscat, dabso, napera, naperb, time
        integer :: nxdiso, nxfree, nseed, imode, nthreads, istat, device, batch, ny, dos_model
        integer :: nmu, imu, off, nb, k, bsz
        integer(8) :: c0, c1, crate
        complex(wp), allocatable :: rmat(:, :, :), tmat(:, :, :)
        integer :: ipos, fpos
        real(wp), allocatable :: tprob(:), rprob(:)
        namelist /waveguide/ boundtype, wol, exfac, dscat, dabso, aratio, nxdiso, nxfree, napera, naperb
        namelist /tpattern/ nseed, imode, nthreads, probatype
        namelist /gpu/ device, precision, batch, ny, dos_model, bintype

        boundtype = "dirichlet"; wol = 4.3_wp; exfac = 1.0_wp; aratio = 1.0_wp; dscat = 1.0_wp; dabso = 0.0_wp
        nxdiso = 100; nxfree = 0; napera = 1.0_wp; naperb = 1.0_wp
        nseed = 1; imode = 1; nthreads = 1; probatype = "incident"
        device = 0; precision = "fp64"; batch = 256; ny = 0; dos_model = 0; bintype = "chebyshev"
        read(unit=funit, nml=waveguide, iostat=istat)
        call rewind_or_stop(funit, istat, .false.)
        read(unit=funit, nml=tpattern, iostat=istat)
        call rewind_or_stop(funit, istat, .false.)
        read(unit=funit, nml=gpu, iostat=istat)
        call rewind_or_stop(funit, istat, .true.)
        if (nseed <= 0) then
            write (error_unit, '(a,i0,a)') "[ERROR] Invalid number of disorder realizations (received ", nseed, &
                "), aborting..."
            stop errmsg_invalid_arg
        else if (probatype /= "incident" .and. probatype /= "total" .and. probatype /= "deposition") then
            write (error_unit, '(a)') "[ERROR] Invalid probatype, received '" // trim(probatype) // &
                "', expected incident/total/deposition, aborting..."
            stop errmsg_invalid_arg
        end if

        call wguide%set_options(ny, dos_model, RG_FP64)
        call wguide%init(boundtype, nxdiso, nxfree, wol, exfac, aratio, dscat, dabso, napera, naperb, &
                         batch_max=max(1, min(batch, nseed)), device=device)
        nmu = wguide%get_nmu()
        if (boundtype == "periodic") then   !! imode=1 is the frontal mode (src/main.f90:772-776)
            imu = mod(imode + nmu/2 - 1, nmu) + 1
        else
            imu = imode
        end if
        ipos = 1                        !! deposition: from the first slice (src/main.f90:769)
        fpos = nxfree + nxdiso/2 + 1    !! to the middle of the medium (:765)
        if (probatype /= "total" .and. (imu < 1 .or. imu > nmu)) then
            write (error_unit, '(2(a,i0),a)') "[ERROR] Invalid incident mode imode=", imode, ", expected in [1, ", nmu, &
                "], aborting..."
            stop errmsg_invalid_arg
        end if
        call wguide%print_parameters(output_unit, "[INFO]")
        write (output_unit, '(2(a,i0),3a)') "[INFO] Nseed=", nseed, ", Imode=", imode, ", Probatype='", &
            trim(probatype), "'"

        bsz = max(1, min(batch, nseed))
        allocate(rmat(nmu, nmu, bsz), tmat(nmu, nmu, bsz), tprob(nmu), rprob(nmu))
        tprob = 0.0_wp
        rprob = 0.0_wp
        call system_clock(c0, crate)
        do off = 0, nseed - 1, bsz
            nb = min(bsz, nseed - off)
            if (probatype == "deposition") then
                call wguide%compute_deposition_batch(1 + off, nb, fpos, ipos, tmat(:, :, 1:nb))
                do k = 1, nb
                    tprob = tprob + abs(tmat(:, imu, k))**2   !! rprob stays zero (:811)
                end do
                cycle
            end if
            call wguide%compute_refl_trans_batch(1 + off, nb, rmat(:, :, 1:nb), tmat(:, :, 1:nb))
            do k = 1, nb
                if (probatype == "incident") then
                    tprob = tprob + abs(tmat(:, imu, k))**2
                    rprob = rprob + abs(rmat(:, imu, k))**2
                else
                    tprob = tprob + sum(abs(tmat(:, :, k))**2, dim=1)
                    rprob = rprob + sum(abs(rmat(:, :, k))**2, dim=1)
                end if
            end do
        end do
        call system_clock(c1)
        time = real(c1 - c0, wp)/real(crate, wp)
        tprob = tprob/nseed
        rprob = rprob/nseed
        write (output_unit, '(a,f0.3,a)') "[INFO] Done, computation time is ", time, " s."
        call save_tpattern(wguide, tprob, rprob, probatype, imode, nseed, time)
        call wguide%del()
    end subroutine

    !! &waveguide and the optional &gpu group, then the plan for passes of up to min(batch, cap)
    !! realizations; returns nxdiso, nxfree and the pass size.
    subroutine open_waveguide(funit, wguide, cap, nxdiso, nxfree, bsz)
        integer, intent(in) :: funit, cap
        type(waveguide_gpu_t), intent(inout) :: wguide
        integer, intent(out) :: nxdiso, nxfree, bsz
        character(len=50) :: boundtype, precision, bintype
        real(wp) :: wol, aratio, exfac, dscat, dabso, napera, naperb
        integer :: istat, device, batch, ny, dos_model
        namelist /waveguide/ boundtype, wol, exfac, dscat, dabso, aratio, nxdiso, nxfree, napera, naperb
        namelist /gpu/ device, precision, batch, ny, dos_model, bintype
        boundtype = "dirichlet"; wol = 4.3_wp; exfac = 1.0_wp; aratio = 1.0_wp; dscat = 1.0_wp; dabso = 0.0_wp
        nxdiso = 100; nxfree = 0; napera = 1.0_wp; naperb = 1.0_wp
        device = 0; precision = "fp64"; batch = 64; ny = 0; dos_model = 0; bintype = "chebyshev"
        read(unit=funit, nml=waveguide, iostat=istat)
        call rewind_or_stop(funit, istat, .false.)
        read(unit=funit, nml=gpu, iostat=istat)
        call rewind_or_stop(funit, istat, .true.)
        bsz = max(1, min(batch, cap))
        call wguide%set_options(ny, dos_model, RG_FP64)
        call wguide%init(boundtype, nxdiso, nxfree, wol, exfac, aratio, dscat, dabso, napera, naperb, &
                         batch_max=bsz, device=device)
        call wguide%print_parameters(output_unit, "[INFO]")
    end subroutine

    !! <dir>/<sub>/result_N<ext>, <dir> the parameter-named folder of generate_directory
    !! (src/waveguide.f90:1046-1067) under "out", with the smallest free N (unique_filename,
    !! src/base_utils.f90:825-840); the directory is created first (create_directory, :846).
    subroutine unique_result(wguide, sub, ext, filename)
        type(waveguide_gpu_t), intent(in) :: wguide
        character(len=*), intent(in) :: sub, ext
        character(len=*), intent(out) :: filename
        character(len=200) :: outputdir
        integer :: ifile
        logical :: fexist
        call wguide%generate_directory("out", outputdir)
        outputdir = trim(outputdir) // sub // "/"
        call execute_command_line("mkdir -p '" // trim(outputdir) // "'")
        do ifile = 1, 1000000
            write (filename, '(a,i0,a)') trim(outputdir) // "result_", ifile, ext
            inquire(file=filename, exist=fexist)
            if (.not. fexist) exit
        end do
    end subroutine

    !! The file of unique_result opened with the timestamp and parameter header (print_timestamp,
    !! print_parameters).
    subroutine open_result(wguide, sub, ext, filename, outfp)
        type(waveguide_gpu_t), intent(in) :: wguide
        character(len=*), intent(in) :: sub, ext
        character(len=*), intent(out) :: filename
        integer, intent(out) :: outfp
        character(len=40) :: zone
        integer :: dt(8)
        call unique_result(wguide, sub, ext, filename)
        open(newunit=outfp, file=filename, action='write')
        call date_and_time(values=dt, zone=zone)
        write (outfp, '(a,i4,a,5(i0.2,a))') "%% Computed on ", dt(1), "-", dt(2), "-", dt(3), " at ", &
            dt(5), ":", dt(6), ":", dt(7), " " // trim(zone) // " by RecurGreen-HIP"
        call wguide%print_parameters(outfp, "%%")
    end subroutine

    !! task_wavefunction and save_wavefunction (src/main.f90:335-380): the field of one realization
    !! in <dir>/wavefun/result_N.wavefun.csv, rows x, y, Re psi, Im psi, Upot.
    subroutine task_wavefunction(funit)
        integer, intent(in) :: funit
        type(waveguide_gpu_t) :: wguide
        integer :: iseed, ipos, imode, istat, nxdiso, nxfree, bsz, nx, ny, ix, iy, outfp
        complex(wp), allocatable :: psi(:, :)
        real(wp), allocatable :: upot(:, :), xmesh(:)
        real(wp) :: u
        character(len=200) :: filename
        namelist /wavefunction/ iseed, ipos, imode
        iseed = 1; ipos = 1; imode = 1
        read(unit=funit, nml=wavefunction, iostat=istat)
        call rewind_or_stop(funit, istat, .false.)
        call open_waveguide(funit, wguide, 1, nxdiso, nxfree, bsz)
        write (output_unit, '(3(a,i0))') "[INFO] Wavefunction parameters: iseed=", iseed, ", ipos=", ipos, ", imode=", imode
        nx = wguide%get_nx()
        ny = wguide%get_ny()
        allocate(psi(ny, nx), upot(nxdiso, ny), xmesh(nx))
        call wguide%compute_wavefunction_field(iseed, ipos, imode, psi)
        call wguide%generate_potential(iseed, upot)
        call wguide%get_xmesh(xmesh)
        call open_result(wguide, "wavefun", ".wavefun.csv", filename, outfp)
        write (outfp, '(3(a,i0))') "%% Wavefunction parameters: iseed=", iseed, ", ipos=", ipos, ", imode=", imode
        write (outfp, '(a)') "x, y, repsi, impsi, upot"
        do ix = 1, nx
            do iy = 1, ny
                u = 0.0_wp
                if (ix - nxfree >= 1 .and. ix - nxfree <= nxdiso) u = upot(ix - nxfree, iy)
                write (outfp, '(g0.16,4(", ",g0.16))') xmesh(ix), (iy - 0.5_wp)/ny, real(psi(iy, ix)), &
                    aimag(psi(iy, ix)), u
            end do
        end do
        close(outfp)
        write (output_unit, '(a)') "[INFO] Result written in file: " // trim(filename)
        call wguide%del()
    end subroutine

    !! task_tprofile and save_tprofile (src/main.f90:402-545): the profiles of realizations 1..nseed
    !! in batched device passes, their average over the realizations with nt > 0, and the raw and
    !! average files under <dir>/tprofile.
    subroutine task_tprofile(funit)
        integer, intent(in) :: funit
        type(waveguide_gpu_t) :: wguide
        real(wp) :: tmin, tmax, time
        integer :: nseed, nthreads, istat, nxdiso, nxfree, bsz, nx, nmu, off, nb, k, v, nsample, outfp, ix
        integer(8) :: c0, c1, crate
        real(wp), allocatable :: tvals(:, :), intens(:, :, :), prof(:, :), avg(:), xmesh(:)
        integer, allocatable :: nt(:)
        character(len=200) :: filename
        character(len=40) :: sfmt
        namelist /tprofile/ tmin, tmax, nseed, nthreads
        tmin = 0.9_wp; tmax = 1.0_wp; nseed = 1; nthreads = 1
        read(unit=funit, nml=tprofile, iostat=istat)
        call rewind_or_stop(funit, istat, .false.)
        if (nseed <= 0) then
            write (error_unit, '(a,i0,a)') "[ERROR] Invalid number of disorder realizations (received ", nseed, &
                "), aborting..."
            stop errmsg_invalid_arg
        else if (tmin < 0.0_wp .or. tmin > 1.0_wp) then
            write (error_unit, '(a,g0.3,a)') "[ERROR] Invalid Tmin, received Tmin=", tmin, ", expected in [0, 1], aborting..."
            stop errmsg_invalid_arg
        else if (tmax < 0.0_wp .or. tmax > 1.0_wp) then
            write (error_unit, '(a,g0.3,a)') "[ERROR] Invalid Tmax, received Tmax=", tmax, ", expected in [0, 1], aborting..."
            stop errmsg_invalid_arg
        else if (tmin >= tmax) then
            write (error_unit, '(a)') "[ERROR] Invalid range (Tmin >= Tmax), aborting..."
            stop errmsg_invalid_arg
        end if
        call open_waveguide(funit, wguide, nseed, nxdiso, nxfree, bsz)
        nx = wguide%get_nx()
        nmu = wguide%get_nmu()
        call wguide%tprofile_reserve(bsz)
        allocate(tvals(nmu, bsz), intens(nmu, nx, bsz), prof(nseed, nx), nt(nseed), avg(nx), xmesh(nx))
        prof = 0.0_wp
        nt = 0
        call system_clock(c0, crate)
        do off = 0, nseed - 1, bsz
            nb = min(bsz, nseed - off)
            call wguide%compute_tprofile_data(1 + off, nb, .false., tvals(:, 1:nb), intens(:, :, 1:nb))
            do k = 1, nb   !! compute_tprofile, step 7 (src/waveguide.f90:887-903)
                do v = 1, nmu
                    if (tmin <= tvals(v, k) .and. tvals(v, k) <= tmax) then
                        prof(off + k, :) = prof(off + k, :) + intens(v, :, k)
                        nt(off + k) = nt(off + k) + 1
                    end if
                end do
                if (nt(off + k) /= 0) prof(off + k, :) = prof(off + k, :)*(4.0_wp/nt(off + k))
            end do
        end do
        call system_clock(c1)
        time = real(c1 - c0, wp)/real(crate, wp)
        nsample = count(nt /= 0)
        avg = 0.0_wp
        if (nsample == 0) then
            write (output_unit, '(a,i0,a)') "[WARN] Found no transmission eigenfunction (nsample=", nsample, &
                "), results will be undetermined..."
        else
            write (output_unit, '(2(a,i0),a)') "[INFO] Found ", nsample, "/", nseed, &
                " transmission eigenfunctions, now averaging..."
            do k = 1, nseed
                if (nt(k) /= 0) avg = avg + prof(k, :)
            end do
            avg = avg/nsample
        end if
        call wguide%get_xmesh(xmesh)
        call open_result(wguide, "tprofile", ".raw.csv", filename, outfp)
        write (outfp, '(3(a,i0),2(a,g0.6),a,f0.3,a)') "%% Transmission profile parameters: Found nsample=", &
            nsample, " / nseed=", nseed, ", nx=", nx, ", tmin=", tmin, ", tmax=", tmax, ", time=", time, " s."
        write (outfp, '(a)') "%% Here follows, on each line, the function |psi_T(x)|^2 for a given realization of the disorder."
        write (sfmt, '(a,i0,a)') "(g0.16,", max(1, nx - 1), "(', ',g0.16))"
        do k = 1, nseed
            write (outfp, sfmt) prof(k, :)
        end do
        close(outfp)
        write (output_unit, '(a)') "[INFO] Raw profile data written in file: " // trim(filename)
        call open_result(wguide, "tprofile", ".avg.csv", filename, outfp)
        write (outfp, '(3(a,i0),2(a,g0.6),a,f0.3,a)') "%% Transmission profile parameters: Found nsample=", &
            nsample, " / nseed=", nseed, ", nx=", nx, ", tmin=", tmin, ", tmax=", tmax, ", time=", time, " s."
        write (outfp, '(a)') "x, tprofile"
        do ix = 1, nx
            write (outfp, '(g0.16,", ",g0.16)') xmesh(ix), avg(ix)
        end do
        close(outfp)
        write (output_unit, '(a)') "[INFO] Average profile written in file: " // trim(filename)
        call wguide%del()
    end subroutine

    !! task_tprofile_group and save_tprofile_group (src/main.f90:564-735): four windows |T - tm| <= dt
    !! of the aperture-filtered matrix, averaged over the realizations with nt > 0, the mean
    !! transmission tavg with its uncertainty, in <dir>/tprofile_group/result_N.csv.
    subroutine task_tprofile_group(funit)
        integer, intent(in) :: funit
        integer, parameter :: nprofile = 4
        type(waveguide_gpu_t) :: wguide
        real(wp) :: tm(nprofile), dt(nprofile), tavg, uctavg, time
        integer :: nseed, nthreads, istat, nxdiso, nxfree, bsz, nx, nv, off, nb, k, v, ip, outfp, ix
        integer :: nsample(nprofile)
        integer(8) :: c0, c1, crate
        real(wp), allocatable :: tvals(:, :), intens(:, :, :), prof(:, :), sumprof(:, :), tavgdata(:), xmesh(:)
        integer :: ntk(nprofile)
        character(len=200) :: filename
        character(len=120) :: sfmt
        namelist /tprofile_group/ tm, dt, nseed, nthreads
        tm = [0.1_wp, 0.4_wp, 0.7_wp, 0.95_wp]; dt = 0.05_wp; nseed = 1; nthreads = 1
        read(unit=funit, nml=tprofile_group, iostat=istat)
        call rewind_or_stop(funit, istat, .false.)
        if (nseed <= 0) then
            write (error_unit, '(a,i0,a)') "[ERROR] Invalid number of disorder realizations (received ", nseed, &
                "), aborting..."
            stop errmsg_invalid_arg
        else if (any(tm < 0.0_wp) .or. any(tm > 1.0_wp)) then
            write (error_unit, '(a)') "[ERROR] Invalid transmission value, expected in [0, 1], aborting..."
            stop errmsg_invalid_arg
        else if (any(dt <= 0.0_wp) .or. any(dt > 0.5_wp)) then
            write (error_unit, '(a)') "[ERROR] Invalid transmission half-width, expected 0 < dt <= 1/2, aborting..."
            stop errmsg_invalid_arg
        end if
        call open_waveguide(funit, wguide, nseed, nxdiso, nxfree, bsz)
        nx = wguide%get_nx()
        nv = wguide%get_neigv_filtered()
        call wguide%tprofile_reserve(bsz)
        allocate(tvals(nv, bsz), intens(nv, nx, bsz), prof(nx, nprofile), sumprof(nx, nprofile), tavgdata(nseed))
        allocate(xmesh(nx))
        sumprof = 0.0_wp
        nsample = 0
        call system_clock(c0, crate)
        do off = 0, nseed - 1, bsz
            nb = min(bsz, nseed - off)
            call wguide%compute_tprofile_data(1 + off, nb, .true., tvals(:, 1:nb), intens(:, :, 1:nb))
            do k = 1, nb   !! compute_tprofile_group, steps 6-7 (src/waveguide.f90:967-990)
                tavgdata(off + k) = sum(tvals(:, k))/nv
                prof = 0.0_wp
                ntk = 0
                do ip = 1, nprofile
                    do v = 1, nv
                        if (abs(tvals(v, k) - tm(ip)) <= dt(ip)) then
                            prof(:, ip) = prof(:, ip) + intens(v, :, k)
                            ntk(ip) = ntk(ip) + 1
                        end if
                    end do
                    if (ntk(ip) /= 0) then
                        sumprof(:, ip) = sumprof(:, ip) + prof(:, ip)*(4.0_wp/ntk(ip))
                        nsample(ip) = nsample(ip) + 1
                    end if
                end do
            end do
        end do
        call system_clock(c1)
        time = real(c1 - c0, wp)/real(crate, wp)
        do ip = 1, nprofile
            if (nsample(ip) /= 0) then
                write (output_unit, '(4(a,i0),a)') "[INFO] Profile ", ip, "/", nprofile, " : Found ", nsample(ip), "/", &
                    nseed, " transmission eigenfunctions, now averaging..."
                sumprof(:, ip) = sumprof(:, ip)/nsample(ip)
            else
                write (output_unit, '(3(a,i0),a)') "[WARN] Profile ", ip, "/", nprofile, &
                    " : Found no transmission eigenstate (nsample=", nsample(ip), "), results will be undetermined..."
            end if
        end do
        tavg = sum(tavgdata)/nseed
        uctavg = sqrt(sum((tavgdata - tavg)**2))/nseed
        call wguide%get_xmesh(xmesh)
        call open_result(wguide, "tprofile_group", ".csv", filename, outfp)
        write (sfmt, '(a,i0,a)') "(a,i0,2(a,g0.6,", nprofile-1, "(', ',g0.6)),a)"
        write (outfp, sfmt) "%% Transmission profile parameters: nx=", nx, ", tm=[", tm, "], dt=[", dt, "]."
        write (sfmt, '(a,i0,a)') "(a,i0,", nprofile-1, "(', ',i0),a,i0,2(a,g0.6),a,f0.3,a)"
        write (outfp, sfmt) "%% Results: Found nsample=[", nsample, "] over nseed=", nseed, &
            ", tavg=", tavg, "+-", uctavg, ", time=", time, " s."
        write (sfmt, '(a,i0,a)') "('x',", nprofile, "(', tprofile',f0.4))"
        write (outfp, sfmt) tm
        write (sfmt, '(a,i0,a)') "(g0.16,", nprofile, "(', ',g0.16))"
        do ix = 1, nx
            write (outfp, sfmt) xmesh(ix), sumprof(ix, :)
        end do
        close(outfp)
        write (output_unit, '(a)') "[INFO] Average profiles written in file: " // trim(filename)
        call wguide%del()
    end subroutine

    !! save_tpattern (src/main.f90:861-901) without the TikZ step.
    subroutine save_tpattern(wguide, tprob, rprob, probatype, imode, nseed, time)
        type(waveguide_gpu_t), intent(in) :: wguide
        real(wp), intent(in) :: tprob(:), rprob(:), time
        character(len=*), intent(in) :: probatype
        integer, intent(in) :: imode, nseed
        character(len=200) :: filename
        character(len=40) :: zone
        integer :: outfp, imu, dt(8)
        call unique_result(wguide, "tpattern", ".csv", filename)
        open(newunit=outfp, file=filename, action='write')
        call date_and_time(values=dt, zone=zone)   !! print_timestamp (src/base_utils.f90:703-716)
        write (outfp, '(a,i4,a,5(i0.2,a))') "%% Computed on ", dt(1), "-", dt(2), "-", dt(3), " at ", &
            dt(5), ":", dt(6), ":", dt(7), " " // trim(zone) // " by RecurGreen-HIP"
        call wguide%print_parameters(outfp, "%%")
        write (outfp, '(2(a,i0),a,f0.3,a)') "%% Radiation pattern parameters: probatype='" // trim(probatype) // &
            "', nseed=", nseed, ", imode=", imode, ", time=", time, " s."
        write (outfp, '(a)') "imu, tprob, rprob"
        do imu = 1, size(tprob)
            write (outfp, '(i0,", ",g0.16,", ",g0.16)') imu, tprob(imu), rprob(imu)
        end do
        close(outfp)
        write (output_unit, '(a)') "[INFO] Average profiles written in file: " // trim(filename)
    end subroutine

    !! save_distrib and save_raw_eigvals (src/main.f90:257-328): the %% header (timestamp, parameters,
    !! Nseed/Nbins/Nthreads/Time/Speed), then "tm, rho" rows, or the eigenvalue count and one value a
    !! line; each file takes the smallest free N of its own name pattern, as unique_filename does.
    subroutine save_files(wguide, rho, tallvals, nseed, nbins, nthreads, time, speed, save_eigvals)
        type(waveguide_gpu_t), intent(in) :: wguide
        real(wp), intent(in) :: rho(:, :), tallvals(:), time, speed
        integer, intent(in) :: nseed, nbins, nthreads
        logical, intent(in) :: save_eigvals
        character(len=200) :: filename
        integer :: outfp, k
        call open_result(wguide, "distrib", ".distrib.csv", filename, outfp)
        write (outfp, '(3(a,i0),2(a,g0.6),a)') "%% Nseed=", nseed, ", Nbins=", nbins, ", Nthreads=", nthreads, &
            ", Time=", time, " s, Speed=", speed, " seed/s"
        write (outfp, '(a)') "tm, rho"
        do k = 1, size(rho, 1)
            write (outfp, '(g0.16,a,g0.16)') rho(k, 1), ", ", rho(k, 2)
        end do
        close(outfp)
        write (output_unit, '(a)') "[INFO] Result written in file: " // trim(filename)
        if (save_eigvals) then
            call open_result(wguide, "distrib", ".eigvals.csv", filename, outfp)
            write (outfp, '(3(a,i0),2(a,g0.6),a)') "%% Nseed=", nseed, ", Nbins=", nbins, ", Nthreads=", nthreads, &
                ", Time=", time, " s, Speed=", speed, " seed/s"
            write (outfp, '(a,i0,a)') "%% Transmission eigenvalues (", size(tallvals), " values)"
            do k = 1, size(tallvals)
                write (outfp, '(g0.16)') tallvals(k)
            end do
            close(outfp)
            write (output_unit, '(a)') "[INFO] Eigenvalues saved in file: " // trim(filename)
        end if
    end subroutine

end program
