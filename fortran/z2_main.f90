! Driver with the structure of the reference program (src/main.f90:1-137):
! init (fresh or from confin.txt), boson warm-up (is_warm), pre, Nbin bins of
! [Wolff global sweep (is_global) + local sweep, with the time-displaced
! measurement after Nthermal bins when is_tau] with the per-bin equal-time and
! time-displaced files appended (src/fourier_trans.f90:251-449), final
! configuration written.
! All chains of seeds.txt (or the first N: second argument, the reference's
! MPI size) run in this one process on the GPU.
! Usage: z2_gauge_gpu.out [fp64|mixed] [N]
program z2_gauge_gpu
    use iso_c_binding
    use DQMC_Model_mod
    use ProcessMatrix
    use Stabilize_mod
    use LocalSweep_mod
    use GlobalUpdate_mod
    use FourierTrans_mod
    implicit none
    integer :: iseed, nbc, nth
    logical :: is_beta, istau_tmp
    integer(kind=8) :: ICPU_1, ICPU_2, N_P_SEC
    real(kind=8) :: xmax, xmean
    type(LocalSweep) :: Sweep_local
    type(GlobalUpdate) :: Sweep_global
    type(FourierTrans) :: Fourier
    type(Propagator), allocatable :: PropU, PropD
    type(WrapList), allocatable :: WrU, WrD
    character(len=32) :: arg

    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg)
        if (trim(arg) == 'mixed') precision_mode = Z2_PREC_MIXED
    endif
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        read(arg, *) max_chains
    endif
    call SYSTEM_CLOCK(COUNT_RATE = N_P_SEC)
    call SYSTEM_CLOCK(COUNT = ICPU_1)
    call Model_init(iseed)
    allocate(PropU); call PropU%make()
    allocate(PropD); call PropD%make()
    allocate(WrU); call WrU%make()
    allocate(WrD); call WrD%make()
    call Stabilize_init()
    call Sweep_local%init()
    if (is_global) call Sweep_global%init()
    open(unit=50, file='info.txt', status='unknown', action='write')
    write(50, *) 'Lx, Ly, Ltrot, Beta, Nwrap, Nbin, chains  :', Nlx, Nly, Ltrot, Beta, Nwrap, Nbin, n_chains
    if (is_warm) then
        do nth = 1, Nwarm
            call Sweep_local%therm(iseed)
        enddo
        call Sweep_local%ctrl_print_t()
    endif
    call Sweep_local%pre(PropU, PropD, WrU, WrD)
    is_beta = .true.; istau_tmp = .false.
    do nbc = 1, Nbin
        if (nbc > Nthermal) istau_tmp = is_tau
        if (is_global) call Sweep_global%sweep(PropU, PropD, WrU, WrD, iseed, is_beta)
        call Sweep_local%sweep(PropU, PropD, WrU, WrD, iseed, is_beta, istau_tmp)
        call write_obs_equal()
        if (istau_tmp) call Fourier%prtau(Obs_tau)
    enddo
    if (is_global) call Sweep_global%ctrl_print()
    call Sweep_local%ctrl_print_l(istau_tmp)
    xmax = PropU%Xmaxm
    xmean = PropU%Xmeanm / dble(2 * Nbin * Nst * Nsweep)
    call SYSTEM_CLOCK(COUNT = ICPU_2)
    write(50, *) 'Max diff Matrix                                :', xmax
    write(50, *) 'Mean diff Matrix                               :', xmean
    write(50, *) 'Tot CPU time                                   :', dble(ICPU_2 - ICPU_1) / dble(N_P_SEC)
    close(50)
    call Sweep_local%clear()
    if (is_global) call Sweep_global%clear()
    call Stabilize_clear()
    call Model_clear(iseed)
contains
    ! m_write_obs_equal, src/fourier_trans.f90:251-278
    subroutine write_obs_equal()
        call append('flux_avg', Obs_equal%flux_avg)
        call append('chi_flux', Obs_equal%chi_flux)
        call append('pair_sc_q0', Obs_equal%pair_sc_q0)
        call append('cdw_pi', Obs_equal%cdw_pi)
        call append('cdw_ratio', Obs_equal%cdw_ratio)
        call append('density', Obs_equal%density)
    end subroutine write_obs_equal
    subroutine append(fname, x)
        character(len=*), intent(in) :: fname
        real(kind=8), intent(in) :: x
        open(unit=80, file=fname, status='unknown', action='write', position='append')
        write(80, *) x
        close(80)
    end subroutine append
end program z2_gauge_gpu
