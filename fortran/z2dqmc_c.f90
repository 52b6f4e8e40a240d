! iso_c_binding interfaces of the gfx950 engine (include/z2dqmc.h).
! Every procedure returns the C status (0 = ok); z2_error_message() returns
! z2_last_error() as a Fortran string.
module z2dqmc_c
    use iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: Z2_PREC_FP64 = 0, Z2_PREC_MIXED = 1
    integer(c_int32_t), parameter :: Z2_BETA_PER_CHAIN = -1

    type, bind(C) :: z2_params
        integer(c_int32_t) :: Nlx, Nly, Ltrot, Nwrap, Nsweep, precision
    end type z2_params

    type, bind(C) :: z2_chain_desc
        real(c_double) :: RT, J, h, mu, Q, Beta
        integer(c_int32_t) :: seed, absolute
    end type z2_chain_desc

    interface
        function z2_last_error() bind(C, name="z2_last_error") result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        function z2_ctx_create(p, n, chains, ctx) bind(C, name="z2_ctx_create") result(rc)
            import :: c_int, c_int32_t, c_ptr, z2_params, z2_chain_desc
            type(z2_params), intent(in) :: p
            integer(c_int32_t), value :: n
            type(z2_chain_desc), intent(in) :: chains(*)
            type(c_ptr), intent(out) :: ctx
            integer(c_int) :: rc
        end function
        function z2_ctx_destroy(ctx) bind(C, name="z2_ctx_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function z2_init_fields(ctx) bind(C, name="z2_init_fields") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function z2_set_sigma(ctx, chain, sigma, seed) bind(C, name="z2_set_sigma") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain, seed
            real(c_double), intent(in) :: sigma(*)
            integer(c_int) :: rc
        end function
        function z2_get_sigma(ctx, chain, sigma) bind(C, name="z2_get_sigma") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            real(c_double), intent(out) :: sigma(*)
            integer(c_int) :: rc
        end function
        function z2_get_seed(ctx, chain, seed) bind(C, name="z2_get_seed") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            integer(c_int32_t), intent(out) :: seed
            integer(c_int) :: rc
        end function
        function z2_pre(ctx) bind(C, name="z2_pre") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function z2_sweep(ctx, is_beta) bind(C, name="z2_sweep") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: is_beta
            integer(c_int) :: rc
        end function
        function z2_wrap_pre(ctx, nt) bind(C, name="z2_wrap_pre") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_wrap_L(ctx, nt, flag) bind(C, name="z2_wrap_L") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt, flag
            integer(c_int) :: rc
        end function
        function z2_wrap_R(ctx, nt, flag) bind(C, name="z2_wrap_R") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt, flag
            integer(c_int) :: rc
        end function
        function z2_slice_L(ctx, nt) bind(C, name="z2_slice_L") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_slice_R(ctx, nt) bind(C, name="z2_slice_R") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_obs_equal_calc(ctx, nt) bind(C, name="z2_obs_equal_calc") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_get_obs_equal(ctx, chain, out) bind(C, name="z2_get_obs_equal") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            real(c_double), intent(out) :: out(7)
            integer(c_int) :: rc
        end function
        function z2_get_drift(ctx, chain, xmax, xmean) bind(C, name="z2_get_drift") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            real(c_double), intent(out) :: xmax, xmean
            integer(c_int) :: rc
        end function
        function z2_get_counters(ctx, chain, out) bind(C, name="z2_get_counters") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            integer(c_int64_t), intent(out) :: out(2)
            integer(c_int) :: rc
        end function
        function z2_get_green(ctx, chain, G) bind(C, name="z2_get_green") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            real(c_double), intent(out) :: G(*)
            integer(c_int) :: rc
        end function
        function z2_enable_tau(ctx) bind(C, name="z2_enable_tau") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function z2_sweep_tau(ctx, is_beta, toggle) bind(C, name="z2_sweep_tau") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: is_beta, toggle
            integer(c_int) :: rc
        end function
        function z2_get_obs_tau(ctx, chain, out) bind(C, name="z2_get_obs_tau") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            real(c_double), intent(out) :: out(*)
            integer(c_int) :: rc
        end function
        function z2_set_lambda(ctx, chain, lam) bind(C, name="z2_set_lambda") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            real(c_double), intent(in) :: lam(*)
            integer(c_int) :: rc
        end function
        function z2_get_lambda(ctx, chain, lam) bind(C, name="z2_get_lambda") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            real(c_double), intent(out) :: lam(*)
            integer(c_int) :: rc
        end function
        function z2_global_sweep(ctx, nglobal) bind(C, name="z2_global_sweep") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nglobal
            integer(c_int) :: rc
        end function
        function z2_get_is_beta(ctx, chain, v) bind(C, name="z2_get_is_beta") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            integer(c_int32_t), intent(out) :: v
            integer(c_int) :: rc
        end function
        function z2_get_global_counters(ctx, chain, out, size_cluster) bind(C, name="z2_get_global_counters") &
                result(rc)
            import :: c_int, c_int32_t, c_ptr, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            integer(c_int64_t), intent(out) :: out(3)
            real(c_double), intent(out) :: size_cluster
            integer(c_int) :: rc
        end function
        function z2_prop_K_pre(ctx, nt) bind(C, name="z2_prop_K_pre") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_prop_T_pre(ctx, nt) bind(C, name="z2_prop_T_pre") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_prop_T_L(ctx, nt) bind(C, name="z2_prop_T_L") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_prop_T_R(ctx, nt) bind(C, name="z2_prop_T_R") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_propgr_K_R(ctx, nt) bind(C, name="z2_propgr_K_R") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_propgr_T_R(ctx, nt) bind(C, name="z2_propgr_T_R") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_localK_prop_L(ctx, nt) bind(C, name="z2_localK_prop_L") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_localK_prop_R(ctx, nt) bind(C, name="z2_localK_prop_R") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_obs_tau_calc(ctx, nt) bind(C, name="z2_obs_tau_calc") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_wrap_tau(ctx, nt) bind(C, name="z2_wrap_tau") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nt
            integer(c_int) :: rc
        end function
        function z2_obs_equal_ave(ctx, nobs) bind(C, name="z2_obs_equal_ave") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nobs
            integer(c_int) :: rc
        end function
        function z2_obs_tau_ave(ctx, nobst) bind(C, name="z2_obs_tau_ave") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nobst
            integer(c_int) :: rc
        end function
        function z2_sweep_reset(ctx, toggle) bind(C, name="z2_sweep_reset") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: toggle
            integer(c_int) :: rc
        end function
        function z2_obs_reset(ctx) bind(C, name="z2_obs_reset") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function z2_dyn_reset(ctx) bind(C, name="z2_dyn_reset") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function z2_obs_tau_reset(ctx) bind(C, name="z2_obs_tau_reset") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function z2_sync(ctx) bind(C, name="z2_sync") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function z2_get_dyn_drift(ctx, chain, xmax, xmean) bind(C, name="z2_get_dyn_drift") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            real(c_double), intent(out) :: xmax(3), xmean(3)
            integer(c_int) :: rc
        end function
        function z2_init_fields_therm(ctx, lx, ly, lt) bind(C, name="z2_init_fields_therm") result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: lx, ly, lt
            integer(c_int) :: rc
        end function
        function z2_set_sigma_therm(ctx, chain, lx, ly, lt, sigma, lam, seed) bind(C, name="z2_set_sigma_therm") &
                result(rc)
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain, lx, ly, lt, seed
            real(c_double), intent(in) :: sigma(*), lam(*)
            integer(c_int) :: rc
        end function
        function z2_therm(ctx) bind(C, name="z2_therm") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function
        function z2_get_therm_counters(ctx, chain, out) bind(C, name="z2_get_therm_counters") result(rc)
            import :: c_int, c_int32_t, c_ptr, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: chain
            integer(c_int64_t), intent(out) :: out(2)
            integer(c_int) :: rc
        end function
        function c_strlen(s) bind(C, name="strlen") result(n)
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: n
        end function
    end interface

contains

    function z2_error_message() result(msg)
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n, i
        p = z2_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        enddo
    end function z2_error_message

    subroutine z2_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc /= 0) then
            write(6, *) "z2dqmc error in ", what, ": ", z2_error_message()
            stop 1
        endif
    end subroutine z2_check

end module z2dqmc_c
