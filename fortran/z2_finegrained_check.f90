! The reference's procedure sequence (Local_sweep_pre / Local_sweep_L / _R with
! Multiply_mod, LocalK_mod, Wrap_pre/L/R, Obs_equal%calc/ave, and with the
! time-displaced measurement Dyn%reset, Dyn%sweep_R -> Obs_tau%calc, propgrK_R,
! propgrT_R, Wrap_tau, Obs_tau%ave) driven through the Fortran modules, one engine
! call per procedure, against the batched z2_pre + z2_sweep_tau on a second
! context of the same chains (paramC_sets.txt, seeds.txt, fresh start).  After
! every bin the fields, seeds, Green's functions and observables of all chains
! must agree bit for bit.  Prints one line per bin; exit status 1 on a mismatch.
! Usage: z2_finegrained_check.out [fp64|mixed]
program z2_finegrained_check
    use iso_c_binding
    use z2dqmc_c
    use DQMC_Model_mod
    use ProcessMatrix
    use Stabilize_mod
    use LocalSweep_mod
    implicit none
    integer :: iseed, nbc, k
    logical :: is_beta, toggle
    type(LocalSweep) :: Sweep_local
    type(Propagator), allocatable :: PropU, PropD
    type(WrapList), allocatable :: WrU, WrD
    type(c_ptr) :: twin
    character(len=32) :: arg
    real(c_double), allocatable :: a(:), b(:)
    integer(c_int32_t) :: sa, sb
    integer(c_int64_t) :: ca(2), cb(2)

    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg)
        if (trim(arg) == 'mixed') precision_mode = Z2_PREC_MIXED
    endif
    call Model_init(iseed)
    allocate(PropU); call PropU%make()
    allocate(PropD); call PropD%make()
    allocate(WrU); call WrU%make()
    allocate(WrD); call WrD%make()
    call Stabilize_init()
    call Sweep_local%init()
    ! the twin: the same chains, advanced by the batched calls
    call z2_check(z2_ctx_create(ctx_params, int(n_chains, c_int32_t), ctx_chains, twin), "z2_ctx_create")
    call z2_check(z2_init_fields(twin), "z2_init_fields")
    if (is_tau) call z2_check(z2_enable_tau(twin), "z2_enable_tau")
    allocate(a(max(2 * Lq * Ltrot, Lq * Lq, 7 * Lq * Ltrot)), b(max(2 * Lq * Ltrot, Lq * Lq, 7 * Lq * Ltrot)))

    fine_grained = .true.
    call Sweep_local%pre(PropU, PropD, WrU, WrD)
    call z2_check(z2_pre(twin), "z2_pre")
    call compare(0, .false.)
    is_beta = .true.
    do nbc = 1, Nbin
        toggle = is_tau .and. nbc > Nthermal
        call Sweep_local%sweep(PropU, PropD, WrU, WrD, iseed, is_beta, toggle)
        call z2_check(z2_sweep_tau(twin, 1_c_int32_t, merge(1_c_int32_t, 0_c_int32_t, toggle)), "z2_sweep_tau")
        call compare(nbc, toggle)
    enddo
    write(6, '(a, i0, a)') 'fine-grained == batched: bit-identical over ', Nbin, ' bins'
    call z2_check(z2_ctx_destroy(twin), "z2_ctx_destroy")
    call Sweep_local%clear()
    call Stabilize_clear()
    call Model_clear(iseed)
contains
    subroutine compare(nbc, toggle)
        integer, intent(in) :: nbc
        logical, intent(in) :: toggle
        integer(c_int32_t) :: c
        integer :: n
        do k = 0, n_chains - 1
            c = int(k, c_int32_t)
            n = 2 * Lq * Ltrot
            call z2_check(z2_get_sigma(ctx, c, a), "z2_get_sigma")
            call z2_check(z2_get_sigma(twin, c, b), "z2_get_sigma")
            call expect(all(a(1:n) == b(1:n)), 'field', nbc)
            call z2_check(z2_get_seed(ctx, c, sa), "z2_get_seed")
            call z2_check(z2_get_seed(twin, c, sb), "z2_get_seed")
            call expect(sa == sb, 'seed', nbc)
            call z2_check(z2_get_counters(ctx, c, ca), "z2_get_counters")
            call z2_check(z2_get_counters(twin, c, cb), "z2_get_counters")
            call expect(all(ca == cb), 'counters', nbc)
            n = Lq * Lq
            call z2_check(z2_get_green(ctx, c, a), "z2_get_green")
            call z2_check(z2_get_green(twin, c, b), "z2_get_green")
            call expect(all(a(1:n) == b(1:n)), 'Green function', nbc)
            if (nbc > 0) then
                call z2_check(z2_get_obs_equal(ctx, c, a), "z2_get_obs_equal")
                call z2_check(z2_get_obs_equal(twin, c, b), "z2_get_obs_equal")
                call expect(all(a(1:7) == b(1:7)), 'equal-time observables', nbc)
            endif
            if (toggle) then
                n = 7 * Lq * Ltrot
                call z2_check(z2_get_obs_tau(ctx, c, a), "z2_get_obs_tau")
                call z2_check(z2_get_obs_tau(twin, c, b), "z2_get_obs_tau")
                call expect(all(a(1:n) == b(1:n)), 'time-displaced observables', nbc)
            endif
        enddo
        write(6, '(a, i0, a, l1, a)') 'bin ', nbc, ' (tau ', toggle, '): identical'
        flush(6)
    end subroutine compare
    subroutine expect(ok, what, nbc)
        logical, intent(in) :: ok
        character(len=*), intent(in) :: what
        integer, intent(in) :: nbc
        if (.not. ok) then
            write(6, '(a, a, a, i0, a, i0)') 'MISMATCH: ', what, ' chain ', k, ' after bin ', nbc
            stop 1
        endif
    end subroutine expect
end program z2_finegrained_check
