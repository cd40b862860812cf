!--------------------------------------------------------------------------------------------------
! pmbd_hip.x geometry.in [settings.in]
!
! Drop-in host for pmbd.x (MBD_rsSCS.f90): reads the same geometry.in / settings.in, prints the
! same result lines, and evaluates MBD@rsSCS on one MI355X through the C ABI of libmbd_hip.so
! (include/mbd_hip.h) instead of MPI + BLACS + ScaLAPACK.
!
!   input semantics     MBD_rsSCS.f90:59-135 (list-directed records, substring keyword match)
!   units               MBD_rsSCS.f90:211-212 (Angstrom -> bohr)
!   printed lines       MBD_rsSCS.f90:141-205, UTILS.f90:223-330, 753-756 (rank-0 lines)
!   timing line         MBD_rsSCS.f90:220-224 (wall time of the whole MBD@rsSCS span)
!
! The solver form is chosen by the environment variable MBD_MODE (fp64 | mixed; default fp64) so
! that an unmodified settings.in stays valid; nprow / npcol are read and ignored (one GPU).
!--------------------------------------------------------------------------------------------------
module mbd_hip_api
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: MBD_MODE_FP64 = 0, MBD_MODE_MIXED = 1

    type, bind(c) :: mbd_diag
        integer(c_int) :: info_scs, fallback_used, refine_iters, n_negative, info_eig
        integer(c_int) :: supercell(3)
        real(c_double) :: residual, t_scs_ms, t_mbd_ms, t_total_ms
    end type mbd_diag

    interface
        integer(c_int) function mbd_create(device, h) bind(c, name="mbd_create")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function mbd_destroy(h) bind(c, name="mbd_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int) function mbd_free_atom_params(sym, c6, alpha, r0) bind(c, name="mbd_free_atom_params")
            import :: c_int, c_char, c_double
            character(kind=c_char), intent(in) :: sym(*)
            real(c_double), intent(out) :: c6, alpha, r0
        end function
        real(c_double) function mbd_beta(flag_xc) bind(c, name="mbd_beta")
            import :: c_int, c_double
            integer(c_int), value :: flag_xc
        end function
        subroutine mbd_casimir_grid(omega, weight) bind(c, name="mbd_casimir_grid")
            import :: c_double
            real(c_double), intent(out) :: omega(21), weight(21)
        end subroutine
        integer(c_int) function mbd_set_system(h, n, xyz, hirsh, alpha0, c6_0, r0, lattice, periodic, &
                                               vacuum, beta, rc_scs, rc_cfdm, rc_sc) bind(c, name="mbd_set_system")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: n, periodic
            real(c_double), intent(in) :: xyz(*), hirsh(*), alpha0(*), c6_0(*), r0(*), lattice(*)
            integer(c_int), intent(in) :: vacuum(3)
            real(c_double), value :: beta, rc_scs, rc_cfdm, rc_sc
        end function
        integer(c_int) function mbd_run(h, mode, e_mbd, alpha_eff, coupled_pol, molpol_eig, c6_eff, &
                                        c6_mol, d) bind(c, name="mbd_run")
            import :: c_int, c_ptr, c_double, mbd_diag
            type(c_ptr), value :: h
            integer(c_int), value :: mode
            real(c_double), intent(out) :: e_mbd, alpha_eff(*), coupled_pol(*), molpol_eig(*), c6_eff(*), c6_mol
            type(mbd_diag), intent(out) :: d
        end function
    end interface
end module mbd_hip_api

program pmbd_hip
    use iso_c_binding
    use mbd_hip_api
    implicit none
    real(8), parameter             :: bohr = 0.52917721d0
    character(len=200)             :: geometry_file, settings_file, string, dster, temp, mode_env
    character(len=500)             :: info_str
    character(len=2), allocatable  :: atom_name(:)
    real(8), allocatable           :: coords(:,:), hirshfeld_volume(:)
    real(8), allocatable           :: a0(:), c60(:), r0(:), alpha_eff(:), c6_eff(:), coupled_pol(:,:)
    real(8)                        :: lattice_vector(3,3), molpol_eig(3,21), omega(21), weight(21)
    real(8)                        :: mbd_supercell_cutoff, mbd_cfdm_dip_cutoff, mbd_scs_dip_cutoff
    real(8)                        :: e_mbd, c6_mol, t_start, t_stop
    logical                        :: mbd_scs_vacuum_axis(3)
    integer                        :: flag_xc, n_periodic, n_atoms, nprow, npcol, i, k, io_file, io_line
    integer                        :: input_settings(7), nargs, rc
    integer(c_int)                 :: vac(3), mode
    integer(8)                     :: clock0, clock1, clock_rate
    type(c_ptr)                    :: h
    type(mbd_diag)                 :: diag

    ! defaults, UTILS.f90:78-93
    flag_xc = 1
    mbd_supercell_cutoff = 25.0d0
    mbd_cfdm_dip_cutoff = 100.0d0
    mbd_scs_dip_cutoff = 120.0d0
    n_periodic = 0
    mbd_scs_vacuum_axis = .false.
    lattice_vector = 0.0d0
    nprow = 1
    npcol = 1
    input_settings = 0
    geometry_file = ""
    settings_file = ""

    nargs = command_argument_count()
    write(*,'(2X,A,I8,1X,A)') "Using  ", 1, "parallel tasks."
    write(*,'(2X,A,I8,1X,A)') "Reading", nargs, "files."
    if (nargs == 0) then
        write(*,*) "ERROR-Missing input file "
        write(*,*) "Please specify input file"
        write(*,*) "... Aborting current run"
        stop
    end if
    call get_command_argument(1, geometry_file)
    if (nargs >= 2) call get_command_argument(2, settings_file)

    ! geometry, MBD_rsSCS.f90:66-89
    open(999, file=trim(geometry_file), status="old", iostat=io_file)
    if (io_file /= 0) then
        write(*,*) "Error in reading input file"
        stop
    end if
    read(999,*) n_atoms
    read(999,*)
    allocate(coords(3,n_atoms), atom_name(n_atoms), hirshfeld_volume(n_atoms))
    do i = 1, n_atoms
        read(999,*,iostat=io_line) atom_name(i), coords(:,i), hirshfeld_volume(i)
    end do
    string = ""
    read(999,*,iostat=io_line) string
    if (index(string, "lattice_vector") >= 1) then
        backspace(999)
        n_periodic = 3
        do i = 1, 3
            read(999,*,iostat=io_line) dster, lattice_vector(:,i)
        end do
    end if
    close(999)
    write(*,*) "lattice vector:"
    write(*,*) lattice_vector

    ! settings, MBD_rsSCS.f90:98-135
    if (nargs >= 2) then
        open(111, file=trim(settings_file), status="old", iostat=io_file)
        if (io_file /= 0) then
            write(*,*) "Error in reading input file"
            stop
        end if
        io_line = 0
        do while (io_line == 0)
            read(111,'(A)',iostat=io_line) string
            if (io_line /= 0) exit
            if (index(string, "xc") >= 1) then
                input_settings(1) = 1
                read(string,*,iostat=io_line) dster, flag_xc
            else if (index(string, "mbd_supercell_cutoff") >= 1) then
                input_settings(2) = 1
                read(string,*,iostat=io_line) dster, mbd_supercell_cutoff
            else if (index(string, "mbd_cfdm_dip_cutoff") >= 1) then
                input_settings(3) = 1
                read(string,*,iostat=io_line) dster, mbd_cfdm_dip_cutoff
            else if (index(string, "mbd_scs_dip_cutoff") >= 1) then
                input_settings(4) = 1
                read(string,*,iostat=io_line) dster, mbd_scs_dip_cutoff
            else if (index(string, "mbd_scs_vacuum_axis") >= 1) then
                input_settings(5) = 1
                read(string,*,iostat=io_line) dster, mbd_scs_vacuum_axis(:)
            else if (index(string, "nprow") >= 1) then
                input_settings(6) = 1
                read(string,*,iostat=io_line) dster, nprow
            else if (index(string, "npcol") >= 1) then
                input_settings(7) = 1
                read(string,*,iostat=io_line) dster, npcol
            end if
        end do
        close(111)
    end if

    ! geometry and settings echo, MBD_rsSCS.f90:141-205
    write(*,'(3x,A)') "|----------------------------------Input Geometry----------------------------------"
    write(*,'(3x,A,I5)') "Number of atoms", n_atoms
    write(*,*)
    do i = 1, n_atoms
        write(*,'(4x,A,4(2x,f15.6))') atom_name(i), coords(:,i), hirshfeld_volume(i)
    end do
    write(*,'(3x,A)') "|----------------------------------------------------------------------------------"
    write(*,'(3x,A)') "|--------------------------------MBD@rSCS Settings---------------------------"
    ! the reference prints the functional twice when xc is given (two leading
    ! blanks before the name) and once, without the blanks, otherwise
    temp = ""
    if (flag_xc == 1) temp = "(PBE)"
    if (flag_xc == 2) temp = "(PBE0)"
    if (flag_xc == 3) temp = "(HSE)"
    if (input_settings(1) == 1) then
        temp = "  "//trim(temp)
        write(*,'(4x,A,I4,A)') "Exchange-Correlation functional:", flag_xc, trim(temp)
    end if
    write(*,'(4x,A)') "xc type not specified in setting file, using default for "
    write(*,'(4x,A,I4,A)') "Exchange-Correlation functional:", flag_xc, trim(temp)
    if (n_periodic > 0) then
        if (input_settings(2) /= 1) write(*,'(4x,A)') "mbd_supercell_cutoff  not specified in setting file, using default for "
        write(*,'(4x,A,f10.5,A)') "Radius used to construct the supercell in periodic MBD calculations", &
            mbd_supercell_cutoff, " Angstrom"
        if (input_settings(3) /= 1) write(*,'(4x,A)') "mbd_cfdm_dip_cutoff  not specified in setting file, using default for "
        write(*,'(4x,A,f10.5,A)') "Radius used to integrate dipole field in periodic MBD calculations", &
            mbd_cfdm_dip_cutoff, " Angstrom"
        if (input_settings(4) /= 1) write(*,'(4x,A)') "mbd_scs_dip_cutoff  not specified in setting file, using default for "
        write(*,'(4x,A,f10.5,A)') "Radius used to integrate dipole field in periodic SCS calculations", &
            mbd_scs_dip_cutoff, " Angstrom"
        if (input_settings(5) == 1) then
            write(*,'(4x,A)') " Found mbd_scs_vacuum_axis keyword in setting file"
            write(*,'(4x,A)') " This keyword specifies direction of vacuum region in case of low-dimensional systems"
            write(*,'(4x,A)') "Default: No vacuum mbd_scs_vacuum_axis .false. .false. .false."
            write(*,'(4x,A)') "For example in the case of periodic slab along xy directions the keyword"
            write(*,'(4x,A)') "mbd_scs_vacuum_axis .false. .false. .true. is needed to specify z as"
            write(*,'(4x,A)') "vacuum direction in MBD/SCS calculation."
        end if
        write(*,'(3x,A)') "|---------------------------------------------------------------------------"
    end if

    ! units, MBD_rsSCS.f90:211-212
    coords = coords / bohr
    lattice_vector = lattice_vector / bohr

    ! free-atom data (an unknown element stops before any device work, UTILS.f90:1693-1699)
    allocate(a0(n_atoms), c60(n_atoms), r0(n_atoms), alpha_eff(n_atoms), c6_eff(n_atoms))
    allocate(coupled_pol(20, n_atoms))
    do i = 1, n_atoms
        if (mbd_free_atom_params(trim(atom_name(i))//c_null_char, c60(i), a0(i), r0(i)) /= 0) then
            write(info_str,'(1X,4A)') '*** WARNING: VdW parameters not defined for atom: ', atom_name(i)
            write(*,*) trim(info_str)
            stop
        end if
    end do
    if (flag_xc < 1 .or. flag_xc > 3) then
        write(*,*) "***  WARNING range seperated parameter beta is not defined for"
        write(*,*) "this fxc defaulting it to 0.0 , which will give MBD energy"
        write(*,*) "WITHOUT short range damping"
    end if

    mode = MBD_MODE_FP64
    call get_environment_variable("MBD_MODE", mode_env)
    if (trim(mode_env) == "mixed") mode = MBD_MODE_MIXED

    rc = mbd_create(0_c_int, h)
    if (rc /= 0) then
        write(*,*) "ERROR: mbd_create failed", rc
        stop 1
    end if
    vac = 0
    do i = 1, 3
        if (mbd_scs_vacuum_axis(i)) vac(i) = 1
    end do
    rc = mbd_set_system(h, int(n_atoms, c_int), coords, hirshfeld_volume, a0, c60, r0, lattice_vector, &
                        int(merge(1, 0, n_periodic > 0), c_int), vac, mbd_beta(int(flag_xc, c_int)), &
                        mbd_scs_dip_cutoff, mbd_cfdm_dip_cutoff, mbd_supercell_cutoff)
    if (rc /= 0) then
        write(*,*) "ERROR: mbd_set_system failed", rc
        stop 1
    end if

    ! the timed span, MBD_rsSCS.f90:220-224
    call system_clock(clock0, clock_rate)
    rc = mbd_run(h, mode, e_mbd, alpha_eff, coupled_pol, molpol_eig, c6_eff, c6_mol, diag)
    call system_clock(clock1)
    if (rc /= 0) then
        write(*,*) "ERROR: mbd_run failed", rc
        stop 1
    end if
    call mbd_casimir_grid(omega, weight)

    ! result lines, UTILS.f90:223-330 (each through WRITE(*,*) TRIM(...))
    call out("(2x,A)", "| Many-Body Dispersion (MBD@rsSCS) energy ")
    call out("(2x,A)", "| System dynamic molecular polarizability alpha(iw) (bohr^3)")
    call out("(2x,A)", "| ---------------------------------------------------------------------------")
    call out("(2x,A)", "|  omega(Ha)   alpha_iso      alpha_xx       alpha_yy       alpha_zz")
    do k = 1, 21
        write(info_str,'(2x,"| ",F10.6,4(e15.6))') omega(k), sum(molpol_eig(:,k))/3.0, &
            molpol_eig(1,k), molpol_eig(2,k), molpol_eig(3,k)
        write(*,*) trim(info_str)
    end do
    call out("(2x,A)", "| ---------------------------------------------------------------------------")
    write(info_str,'(2x,A,f20.8,2x,A)') "| System C6 coefficient    :  ", c6_mol, "  hartree bohr^6"
    write(*,*) trim(info_str)
    call out("(2x,A)", "| ---------------------------------------------------------------------------")
    call out("(2x,A)", "| Partitioned C6 (hartree bohr^6) | alpha (bohr^3)  of atom in system")
    call out("(2x,A)", "| ---------------------------------------------------------------------------")
    do i = 1, n_atoms
        write(info_str,'(2x,A,I4,2x,A,f12.6,6x,f12.6)') "| ATOM ", i, atom_name(i), c6_eff(i), alpha_eff(i)
        write(*,*) trim(info_str)
    end do
    call out("(2x,A)", "| ---------------------------------------------------------------------------")
    write(info_str,'(2X,A,i3,A,i3,A,i3,A)') "| Creating super cell of dimension", diag%supercell(1), " X ", &
        diag%supercell(2), " X ", diag%supercell(3), " in MBD calculation"
    write(*,*) trim(info_str)
    write(info_str,'(2x,"| containing", I6,A)') n_atoms*diag%supercell(1)*diag%supercell(2)*diag%supercell(3), "  atom"
    write(*,*) trim(info_str)
    if (diag%n_negative > 0) then
        write(info_str,'(A,I4,A)') "***WARNING: found ", diag%n_negative, " negative eigenvalues in MBD energy calculation."
        write(*,*) trim(info_str)
    end if
    write(info_str,'(2X,A,F20.9,A,F20.9,A)') "| MBD@rsSCS energy   :", e_mbd, " Ha     ", e_mbd*27.211, " eV"
    write(*,*) trim(info_str)
    t_start = 0.0d0
    t_stop = dble(clock1 - clock0) / dble(clock_rate)
    write(*,'(A,F18.10)') "Total Calculation time: ", t_stop - t_start
    rc = mbd_destroy(h)

contains

    subroutine out(fmt, text)
        character(len=*), intent(in) :: fmt, text
        character(len=500) :: line
        write(line, fmt) text
        write(*,*) trim(line)
    end subroutine out

end program pmbd_hip
