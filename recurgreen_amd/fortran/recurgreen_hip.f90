!!**************************************************************************************************
!! recurgreen_hip.f90 -- Fortran binding of the gfx950 RGF library (include/recurgreen_hip.h).
!!
!! Drop-in for the CPU/LAPACK transmission path of RecurGreen's `waveguide` module: the type
!! waveguide_gpu_t keeps the argument lists of waveguide_t%init (src/waveguide.f90:81),
!! get_nmu/get_neigv/get_nx (:291-314), compute_tdistrib (:392) and del (:275), and maps every
!! non-zero library status to `stop` with the reference's message, as the reference does on
!! invalid arguments (:91-140) and LAPACK failures (src/base_utils.f90:255-269).
!!**************************************************************************************************
module recurgreen_hip
    use, intrinsic :: iso_c_binding
    use, intrinsic :: iso_fortran_env, only: real64, int64, error_unit
    implicit none
    private

    integer, parameter, public :: wp = real64
    integer(c_int32_t), parameter, public :: RG_FP64 = 0, RG_MIXED_FP32 = 1, RG_MIXED_BF16X3 = 2
    character(len=*), parameter, public :: errmsg_invalid_arg = "RecurGreen: Invalid arguments, aborting..."

    !! struct rg_params (include/recurgreen_hip.h)
    type, bind(C), public :: rg_params_t
        integer(c_int32_t) :: boundtype
        integer(c_int32_t) :: nxdiso, nxfree
        real(c_double) :: wol, exfac, aratio, dscat, dabso, napera, naperb
        integer(c_int32_t) :: ny
        integer(c_int32_t) :: dos_model
    end type

    interface
        integer(c_int) function rg_plan_create(params, batch_max, device, plan) bind(C, name="rg_plan_create")
            import :: c_int, c_int32_t, c_ptr, rg_params_t
            type(rg_params_t), intent(in) :: params
            integer(c_int32_t), value :: batch_max, device
            type(c_ptr), intent(out) :: plan
        end function
        integer(c_int) function rg_plan_destroy(plan) bind(C, name="rg_plan_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: plan
        end function
        integer(c_int) function rg_plan_set_option(plan, key, value) bind(C, name="rg_plan_set_option")
            import :: c_int, c_int32_t, c_ptr, c_char
            type(c_ptr), value :: plan
            character(kind=c_char), dimension(*), intent(in) :: key
            integer(c_int32_t), value :: value
        end function
        integer(c_int) function rg_plan_dims(plan, dims, vals) bind(C, name="rg_plan_dims")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: plan
            integer(c_int32_t), intent(out) :: dims(11)
            real(c_double), intent(out) :: vals(4)
        end function
        integer(c_int) function rg_tdistrib(plan, seed0, nseed, precision, tallvals, status) bind(C, name="rg_tdistrib")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: plan
            integer(c_int64_t), value :: seed0
            integer(c_int32_t), value :: nseed, precision
            real(c_double), intent(out) :: tallvals(*)
            integer(c_int32_t), intent(out) :: status(*)
        end function
        integer(c_int) function rg_refl_trans(plan, seed0, nseed, rmat, tmat, status) bind(C, name="rg_refl_trans")
            import :: c_int, c_int32_t, c_int64_t, c_double_complex, c_ptr
            type(c_ptr), value :: plan
            integer(c_int64_t), value :: seed0
            integer(c_int32_t), value :: nseed
            complex(c_double_complex), intent(out) :: rmat(*), tmat(*)
            integer(c_int32_t), intent(out) :: status(*)
        end function
        integer(c_int) function rg_deposition(plan, seed0, nseed, fpos, ipos, depmat, status) bind(C, name="rg_deposition")
            import :: c_int, c_int32_t, c_int64_t, c_double_complex, c_ptr
            type(c_ptr), value :: plan
            integer(c_int64_t), value :: seed0
            integer(c_int32_t), value :: nseed, fpos, ipos
            complex(c_double_complex), intent(out) :: depmat(*)
            integer(c_int32_t), intent(out) :: status(*)
        end function
        integer(c_int) function rg_tprofile_reserve(plan, nreal_max) bind(C, name="rg_tprofile_reserve")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: plan
            integer(c_int32_t), value :: nreal_max
        end function
        integer(c_int) function rg_tprofile(plan, seed0, nseed, filtered, tvals, intensity, status) &
                bind(C, name="rg_tprofile")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: plan
            integer(c_int64_t), value :: seed0
            integer(c_int32_t), value :: nseed, filtered
            real(c_double), intent(out) :: tvals(*), intensity(*)
            integer(c_int32_t), intent(out) :: status(*)
        end function
        integer(c_int) function rg_wavefunction(plan, seed, ipos, imode, psi, status) bind(C, name="rg_wavefunction")
            import :: c_int, c_int32_t, c_int64_t, c_double_complex, c_ptr
            type(c_ptr), value :: plan
            integer(c_int64_t), value :: seed
            integer(c_int32_t), value :: ipos, imode
            complex(c_double_complex), intent(out) :: psi(*)
            integer(c_int32_t), intent(out) :: status(*)
        end function
        integer(c_int) function rg_generate_potential(plan, seed0, nreal, upot) bind(C, name="rg_generate_potential")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: plan
            integer(c_int64_t), value :: seed0
            integer(c_int32_t), value :: nreal
            real(c_double), intent(out) :: upot(*)
        end function
        integer(c_int) function rg_histogram(plan, samples, nsamp, on_device, bins, nbins, counts, missing, dropped) &
                bind(C, name="rg_histogram")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            type(c_ptr), value :: plan
            real(c_double), intent(in) :: samples(*)
            integer(c_int64_t), value :: nsamp
            integer(c_int32_t), value :: on_device
            real(c_double), intent(in) :: bins(*)
            integer(c_int32_t), value :: nbins
            integer(c_int64_t), intent(out) :: counts(*), missing, dropped
        end function
        type(c_ptr) function rg_status_string(code) bind(C, name="rg_status_string")
            import :: c_ptr, c_int32_t
            integer(c_int32_t), value :: code
        end function
    end interface

    !! Counterpart of waveguide_t (src/waveguide.f90:13-73) whose sweep runs on the GPU.
    type, public :: waveguide_gpu_t
        private
        type(c_ptr) :: plan = c_null_ptr
        type(rg_params_t) :: params
        integer :: nmu = 0, ny = 0, nx = 0, neigv = 0, nprop = 0
        integer :: imin = 0, imax = 0, jmin = 0, jmax = 0
        character(len=50) :: boundtype = ""
        integer(c_int32_t) :: precision = RG_FP64
    contains
        procedure :: init => init_waveguide_gpu
        procedure :: set_options => set_options_waveguide_gpu
        procedure :: set_plan_option => set_plan_option_waveguide_gpu
        procedure :: get_nmu => get_nmu_waveguide_gpu
        procedure :: get_neigv => get_neigv_waveguide_gpu
        procedure :: get_nx => get_nx_waveguide_gpu
        procedure :: compute_tdistrib => compute_tdistrib_waveguide_gpu
        procedure :: compute_tdistrib_batch => compute_tdistrib_batch_waveguide_gpu
        procedure :: compute_refl_trans => compute_refl_trans_waveguide_gpu
        procedure :: compute_refl_trans_batch => compute_refl_trans_batch_waveguide_gpu
        procedure :: compute_deposition_batch => compute_deposition_batch_waveguide_gpu
        procedure :: tprofile_reserve => tprofile_reserve_waveguide_gpu
        procedure :: compute_tprofile_data => compute_tprofile_data_waveguide_gpu
        procedure :: compute_wavefunction_field => compute_wavefunction_field_waveguide_gpu
        procedure :: generate_potential => generate_potential_waveguide_gpu
        procedure :: get_xmesh => get_xmesh_waveguide_gpu
        procedure :: get_ny => get_ny_waveguide_gpu
        procedure :: get_neigv_filtered => get_neigv_filtered_waveguide_gpu
        procedure :: histogram => histogram_waveguide_gpu
        procedure :: print_parameters => print_parameters_waveguide_gpu
        procedure :: generate_directory => generate_directory_waveguide_gpu
        procedure :: del => del_waveguide_gpu
    end type

    public :: status_text

contains

    function status_text(code) result(txt)
        integer(c_int), intent(in) :: code
        character(len=:), allocatable :: txt
        character(kind=c_char), pointer :: s(:)
        integer :: n
        call c_f_pointer(rg_status_string(int(code, c_int32_t)), s, [256])
        n = 0
        do while (s(n+1) /= c_null_char .and. n < 255)
            n = n + 1
        end do
        allocate(character(len=n) :: txt)
        txt = transfer(s(1:n), txt)
    end function

    subroutine check_status(code)
        integer(c_int), intent(in) :: code
        if (code /= 0) then
            write (error_unit, '(a)') "[ERROR] " // status_text(code)
            stop 1
        end if
    end subroutine

    !! Extra settings with no counterpart in the reference (ny > nmu config mode, DOS model,
    !! precision, device, batch size); must be called before init.
    !! A launch option of the plan (rg_plan_set_option, keys in include/recurgreen_hip.h), after init.
    subroutine set_plan_option_waveguide_gpu(self, key, value)
        class(waveguide_gpu_t), intent(inout) :: self
        character(len=*), intent(in) :: key
        integer, intent(in) :: value
        call check_status(rg_plan_set_option(self%plan, trim(key) // c_null_char, int(value, c_int32_t)))
    end subroutine

    subroutine set_options_waveguide_gpu(self, ny, dos_model, precision)
        class(waveguide_gpu_t), intent(inout) :: self
        integer, intent(in) :: ny, dos_model, precision
        self%params%ny = int(ny, c_int32_t)
        self%params%dos_model = int(dos_model, c_int32_t)
        self%precision = int(precision, c_int32_t)
    end subroutine

    !! waveguide_t%init (src/waveguide.f90:81) with a device plan holding `batch_max` realizations.
    subroutine init_waveguide_gpu(self, boundtype, nxdiso, nxfree, wol, exfac, aratio, dscat, dabso, napera, naperb, &
                                  batch_max, device)
        class(waveguide_gpu_t), intent(inout) :: self
        character(len=*), intent(in) :: boundtype
        integer, intent(in) :: nxdiso, nxfree
        real(wp), intent(in) :: wol, exfac, aratio, dscat, dabso, napera, naperb
        integer, intent(in), optional :: batch_max, device
        integer(c_int32_t) :: dims(11), bmax, dev
        real(c_double) :: vals(4)
        select case (boundtype)
            case ("dirichlet")
                self%params%boundtype = 0
            case ("periodic")
                self%params%boundtype = 1
            case default
                write (error_unit, '(a)') "[ERROR] Unknown boundary condition type '" // trim(boundtype) // "', aborting..."
                stop errmsg_invalid_arg
        end select
        self%boundtype = boundtype
        self%params%nxdiso = int(nxdiso, c_int32_t)
        self%params%nxfree = int(nxfree, c_int32_t)
        self%params%wol = wol
        self%params%exfac = exfac
        self%params%aratio = aratio
        self%params%dscat = dscat
        self%params%dabso = dabso
        self%params%napera = napera
        self%params%naperb = naperb
        bmax = 1024
        dev = 0
        if (present(batch_max)) bmax = int(batch_max, c_int32_t)
        if (present(device)) dev = int(device, c_int32_t)
        call check_status(rg_plan_create(self%params, bmax, dev, self%plan))
        call check_status(rg_plan_dims(self%plan, dims, vals))
        self%nmu = dims(1); self%ny = dims(2); self%nx = dims(3); self%neigv = dims(4)
        self%imin = dims(5); self%imax = dims(6); self%jmin = dims(7); self%jmax = dims(8); self%nprop = dims(9)
    end subroutine

    integer function get_nmu_waveguide_gpu(self) result(nmu)
        class(waveguide_gpu_t), intent(in) :: self
        nmu = self%nmu
    end function

    integer function get_neigv_waveguide_gpu(self) result(neigv)
        class(waveguide_gpu_t), intent(in) :: self
        neigv = self%neigv
    end function

    integer function get_nx_waveguide_gpu(self) result(nx)
        class(waveguide_gpu_t), intent(in) :: self
        nx = self%nx
    end function

    !! compute_tdistrib (src/waveguide.f90:392-462): chunk `iseed` of tallvals, same checks (:407-421).
    subroutine compute_tdistrib_waveguide_gpu(self, iseed, tallvals)
        class(waveguide_gpu_t), intent(inout) :: self
        integer, intent(in) :: iseed
        real(wp), intent(inout) :: tallvals(:)
        integer :: nchunk
        integer(c_int32_t) :: st(1)
        real(c_double), allocatable :: t(:)
        nchunk = size(tallvals)/self%neigv
        if (nchunk == 0) then
            write (error_unit, '(a,i0,a,i0,a)') "[ERROR] Invalid eigenvalue array length, received ", size(tallvals), &
                ", expected at least ", self%neigv, "..."
            stop errmsg_invalid_arg
        else if (nchunk*self%neigv /= size(tallvals)) then
            write (error_unit, '(a,i0,a,i0,a)') "[ERROR] Invalid eigenvalue array length, received ", size(tallvals), &
                ", expected ", nchunk*self%neigv, "..."
            stop errmsg_invalid_arg
        else if (iseed < 1 .or. nchunk < iseed) then
            write (error_unit, '(a,i0,a,i0,a)') "[ERROR] Invalid chunk index, received ", iseed, ", expected in [1, ", &
                nchunk, "]..."
            stop errmsg_invalid_arg
        end if
        allocate(t(self%neigv))
        call check_status(rg_tdistrib(self%plan, int(iseed, c_int64_t), 1_c_int32_t, self%precision, t, st))
        tallvals((iseed-1)*self%neigv+1 : iseed*self%neigv) = t
    end subroutine

    !! The whole OpenMP loop of src/main.f90:153-166 as one device pass: seeds 1..nseed.
    subroutine compute_tdistrib_batch_waveguide_gpu(self, seed0, nseed, tallvals)
        class(waveguide_gpu_t), intent(inout) :: self
        integer, intent(in) :: seed0, nseed
        real(wp), intent(inout) :: tallvals(:)
        integer(c_int32_t), allocatable :: st(:)
        if (size(tallvals) /= nseed*self%neigv) then
            write (error_unit, '(a)') "[ERROR] Invalid eigenvalue array length."
            stop errmsg_invalid_arg
        end if
        allocate(st(nseed))
        call check_status(rg_tdistrib(self%plan, int(seed0, c_int64_t), int(nseed, c_int32_t), self%precision, tallvals, st))
    end subroutine

    !! compute_refl_trans_v1 (src/waveguide.f90:469-519) of realization iseed: rmat and tmat are
    !! nmu x nmu, r = -I + 2i s G11 s and t = 2i s G(L;0) s with s = sqrt(Re mu) (:508-512).
    subroutine compute_refl_trans_waveguide_gpu(self, iseed, rmat, tmat)
        class(waveguide_gpu_t), intent(inout) :: self
        integer, intent(in) :: iseed
        complex(wp), intent(inout), contiguous :: rmat(:, :), tmat(:, :)
        integer(c_int32_t) :: st(1)
        call check_refl_trans_shape(self, size(rmat, 1), size(rmat, 2), 1, shape(tmat), 1)
        call check_status(rg_refl_trans(self%plan, int(iseed, c_int64_t), 1_c_int32_t, rmat, tmat, st))
    end subroutine

    !! Realizations seed0 .. seed0+nseed-1 in batched device passes: rmat(:, :, k) and tmat(:, :, k)
    !! belong to realization seed0+k-1.
    subroutine compute_refl_trans_batch_waveguide_gpu(self, seed0, nseed, rmat, tmat)
        class(waveguide_gpu_t), intent(inout) :: self
        integer, intent(in) :: seed0, nseed
        complex(wp), intent(inout), contiguous :: rmat(:, :, :), tmat(:, :, :)
        integer(c_int32_t), allocatable :: st(:)
        call check_refl_trans_shape(self, size(rmat, 1), size(rmat, 2), size(rmat, 3), shape(tmat), nseed)
        allocate(st(nseed))
        call check_status(rg_refl_trans(self%plan, int(seed0, c_int64_t), int(nseed, c_int32_t), rmat, tmat, st))
    end subroutine

    !! compute_deposition (src/waveguide.f90:710-745) for realizations seed0 .. seed0+nseed-1:
    !! depmat(:, :, k) = 2i s G(fpos; ipos) s of realization seed0+k-1, with the position checks of
    !! :719-727.
    subroutine compute_deposition_batch_waveguide_gpu(self, seed0, nseed, fpos, ipos, depmat)
        class(waveguide_gpu_t), intent(inout) :: self
        integer, intent(in) :: seed0, nseed, fpos, ipos
        complex(wp), intent(inout), contiguous :: depmat(:, :, :)
        integer(c_int32_t), allocatable :: st(:)
        if (fpos < 1 .or. self%nx < fpos) then
            write (error_unit, '(2(a,i0),a)') "[ERROR] Invalid final position index, received fpos=", fpos, &
                ", expected in 1..", self%nx, ", aborting..."
            stop errmsg_invalid_arg
        else if (ipos < 1 .or. self%nx < ipos) then
            write (error_unit, '(2(a,i0),a)') "[ERROR] Invalid initial position index, received ipos=", ipos, &
                ", expected in 1..", self%nx, ", aborting..."
            stop errmsg_invalid_arg
        end if
        call check_refl_trans_shape(self, size(depmat, 1), size(depmat, 2), size(depmat, 3), shape(depmat), nseed)
        allocate(st(nseed))
        call check_status(rg_deposition(self%plan, int(seed0, c_int64_t), int(nseed, c_int32_t), &
            int(fpos, c_int32_t), int(ipos, c_int32_t), depmat, st))
    end subroutine

    !! Workspace of the tprofile and wavefunction passes (the Green functions of every slice for
    !! nreal_max realizations per device pass).
    subroutine tprofile_reserve_waveguide_gpu(self, nreal_max)
        class(waveguide_gpu_t), intent(inout) :: self
        integer, intent(in) :: nreal_max
        call check_status(rg_tprofile_reserve(self%plan, int(nreal_max, c_int32_t)))
    end subroutine

    !! Number of transmission eigenvalues of the aperture-filtered matrix (jmax-jmin+1).
    integer function get_neigv_filtered_waveguide_gpu(self) result(nv)
        class(waveguide_gpu_t), intent(in) :: self
        nv = self%jmax - self%jmin + 1
    end function

    !! The data of compute_tprofile / compute_tprofile_group (src/waveguide.f90:843-997) for
    !! realizations seed0 .. seed0+nseed-1: tvals(:, k) the eigenvalues of t^H t (descending),
    !! intensity(:, ix, k) = |G(x_ix;1) s U(:,v)|^2 per eigenvector v; filtered: t over the apertures.
    subroutine compute_tprofile_data_waveguide_gpu(self, seed0, nseed, filtered, tvals, intensity)
        class(waveguide_gpu_t), intent(inout) :: self
        integer, intent(in) :: seed0, nseed
        logical, intent(in) :: filtered
        real(wp), intent(inout), contiguous :: tvals(:, :), intensity(:, :, :)
        integer(c_int32_t), allocatable :: st(:)
        integer(c_int32_t) :: f
        f = 0
        if (filtered) f = 1
        allocate(st(nseed))
        call check_status(rg_tprofile(self%plan, int(seed0, c_int64_t), int(nseed, c_int32_t), f, tvals, intensity, st))
    end subroutine

    !! The field of compute_wavefunction (src/waveguide.f90:636-697): psi(iy, ix) = psi(x_ix, y_iy).
    subroutine compute_wavefunction_field_waveguide_gpu(self, iseed, ipos, imode, psi)
        class(waveguide_gpu_t), intent(inout) :: self
        integer, intent(in) :: iseed, ipos, imode
        complex(wp), intent(inout), contiguous :: psi(:, :)
        integer(c_int32_t) :: st(1)
        if (imode < 1 .or. self%nmu < imode) then
            write (error_unit, '(3(a,i0),a)') "[ERROR] Invalid initial mode index, received imode=", imode, &
                ", expected in 1..", self%nmu, ", aborting..."
            stop errmsg_invalid_arg
        else if (ipos < 1 .or. self%nx < ipos) then
            write (error_unit, '(2(a,i0),a)') "[ERROR] Invalid initial position index, received ipos=", ipos, &
                ", expected in 1..", self%nx, ", aborting..."
            stop errmsg_invalid_arg
        end if
        call check_status(rg_tprofile_reserve(self%plan, 1_c_int32_t))
        call check_status(rg_wavefunction(self%plan, int(iseed, c_int64_t), int(ipos, c_int32_t), int(imode, c_int32_t), &
            psi, st))
    end subroutine

    !! The disorder upot(ix, iy) = U dx/k of realization `seed` (src/waveguide.f90:366-372 layout).
    subroutine generate_potential_waveguide_gpu(self, seed, upot)
        class(waveguide_gpu_t), intent(inout) :: self
        integer, intent(in) :: seed
        real(wp), intent(inout), contiguous :: upot(:, :)
        call check_status(rg_generate_potential(self%plan, int(seed, c_int64_t), 1_c_int32_t, upot))
    end subroutine

    integer function get_ny_waveguide_gpu(self) result(ny)
        class(waveguide_gpu_t), intent(in) :: self
        ny = self%ny
    end function

    !! get_xmesh (src/waveguide.f90:319-323): x_i/L = aratio (i - 1/2 - nxfree)/nxdiso.
    subroutine get_xmesh_waveguide_gpu(self, xmesh)
        class(waveguide_gpu_t), intent(in) :: self
        real(wp), intent(out) :: xmesh(:)
        integer :: i
        do i = 1, self%nx
            xmesh(i) = self%params%aratio*(i - 0.5_wp - self%params%nxfree)/self%params%nxdiso
        end do
    end subroutine

    subroutine check_refl_trans_shape(self, n1, n2, n3, tshape, nseed)
        class(waveguide_gpu_t), intent(in) :: self
        integer, intent(in) :: n1, n2, n3, tshape(:), nseed
        integer :: t3
        t3 = 1
        if (size(tshape) == 3) t3 = tshape(3)
        if (n1 /= self%nmu .or. n2 /= self%nmu .or. tshape(1) /= self%nmu .or. tshape(2) /= self%nmu &
                .or. n3 < nseed .or. t3 < nseed .or. nseed < 1) then
            write (error_unit, '(2(a,i0),a)') "[ERROR] Invalid reflection/transmission array shape, expected ", self%nmu, &
                " x ", self%nmu, " per realization..."
            stop errmsg_invalid_arg
        end if
    end subroutine

    !! Device histogram with the rules of src/base_utils.f90:491-516 (counts returned as reals).
    subroutine histogram_waveguide_gpu(self, samples, bins, histo, missing)
        class(waveguide_gpu_t), intent(inout) :: self
        real(wp), intent(in) :: samples(:), bins(:)
        real(wp), intent(out) :: histo(:), missing
        integer(c_int64_t), allocatable :: counts(:)
        integer(c_int64_t) :: miss, drop
        allocate(counts(size(bins)-1))
        call check_status(rg_histogram(self%plan, samples, int(size(samples), c_int64_t), 0_c_int32_t, bins, &
            int(size(bins)-1, c_int32_t), counts, miss, drop))
        histo = real(counts, wp)
        missing = real(miss, wp)
    end subroutine

    !! print_parameters_waveguide (src/waveguide.f90:1003-1018), same three lines.
    subroutine print_parameters_waveguide_gpu(self, fp, prefix)
        class(waveguide_gpu_t), intent(in) :: self
        integer, intent(in) :: fp
        character(len=*), intent(in) :: prefix
        write (fp, '(2(a,f0.6),2(a,i0))') &
            trim(prefix) // " Waveguide 2D with boundtype='" // trim(self%boundtype) // "', W/lambda=", self%params%wol, &
            ", exfac=", self%params%exfac, ", Nmode=", self%nmu, ", Nprop=", self%nprop
        write (fp, '(2(a,f0.6,a,i0),a,i0)') &
            trim(prefix) // " Filtering: NAper_a=", self%params%napera, " (Nmode_a=", (self%jmax-self%jmin+1), &
            "), NAper_b=", self%params%naperb, " (Nmode_b=", (self%imax-self%imin+1), "), Neigv=", self%neigv
        write (fp, '(2(a,f0.6),2(a,i0),a,f0.6)') &
            trim(prefix) // " Medium: L/lscat=", self%params%dscat, ", L/labso=", self%params%dabso, &
            ", Nx_diso=", self%params%nxdiso, ", Nx_free=", self%params%nxfree, ", aratio=L/W=", self%params%aratio
    end subroutine

    !! The result folder of generate_directory_waveguide (src/waveguide.f90:1046-1067):
    !! <root>/dscat_<L/ls>/dabso_<L/la>/wol_<W/lambda>/<boundtype>/, each number written with six
    !! decimals, a zero before the point below one, and the trailing zeros dropped ("dabso_0.").
    subroutine generate_directory_waveguide_gpu(self, root, outputdir)
        class(waveguide_gpu_t), intent(in) :: self
        character(len=*), intent(in) :: root
        character(len=*), intent(out) :: outputdir
        outputdir = trim(root) // "/" // folder_value("dscat_", self%params%dscat) // "/" // &
            folder_value("dabso_", self%params%dabso) // "/" // folder_value("wol_", self%params%wol) // &
            "/" // trim(self%boundtype) // "/"
    end subroutine

    !! "<name><value>" with the value written as above.
    function folder_value(name, x) result(res)
        character(len=*), intent(in) :: name
        real(c_double), intent(in) :: x
        character(len=:), allocatable :: res
        character(len=40) :: buf
        integer :: n
        write (buf, '(f0.6)') x
        buf = adjustl(buf)
        if (buf(1:1) == ".") buf = "0" // buf
        n = len_trim(buf)
        do while (n > 1 .and. buf(n:n) == "0")
            n = n - 1
        end do
        res = name // buf(1:n)
    end function

    subroutine del_waveguide_gpu(self)
        class(waveguide_gpu_t), intent(inout) :: self
        integer(c_int) :: st
        if (c_associated(self%plan)) then
            st = rg_plan_destroy(self%plan)
            self%plan = c_null_ptr
        end if
    end subroutine

end module
