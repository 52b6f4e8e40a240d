! Fortran API surface of the reference (module and procedure names of
! src/model.f90, src/process_matrix.f90, src/multiply.f90, src/localK.f90,
! src/stabilization.f90, src/obser_equal.f90, src/obser_tau.f90,
! src/dynamics.f90, src/local_sweep.f90) over the gfx950 engine.  Propagator,
! PropGreen and WrapList live on the GPU; the Fortran types keep the
! reference's argument lists.  The two spin copies go through identical
! operations in the reference (src/localK.f90:85-118, src/multiply.f90:8-130),
! so the engine keeps one copy: procedures called for a spin-down object (the
! second object of each type made, as the reference makes PropU then PropD)
! do nothing, and pair procedures (PropU, PropD) act once.
module DQMC_Model_mod
    use iso_c_binding
    use z2dqmc_c
    implicit none
    ! paramC_sets.txt (src/calc_basic.f90:12-44)
    real(kind=8) :: RT, J, h, mu, Q, Beta, Dtau
    integer :: Nlx, Nly, Ltrot, NlxTherm, NlyTherm, LtrotTherm
    integer :: Nwrap, Nbin, Nsweep, Nthermal, Nwarm, Nglobal, absolute, Lq, Ndim, Nst
    logical :: is_tau, is_warm, is_global
    integer :: precision_mode = 0
    ! all chains of this process (one reference MPI rank each); max_chains > 0
    ! takes only the first max_chains seeds (the reference's MPI size)
    integer :: n_chains = 1
    integer :: max_chains = 0
    type(c_ptr) :: ctx = c_null_ptr
    ! the parameters and chains the context was made from (for a second context of
    ! the same chains)
    type(z2_params) :: ctx_params
    type(z2_chain_desc), allocatable :: ctx_chains(:)
contains
    ! read_input + Params_set (src/calc_basic.f90:50-102)
    subroutine read_input()
        open(unit=20, file='paramC_sets.txt', status='old')
        read(20, *) RT, J, h, mu, Q
        read(20, *) Nlx, Nly, Ltrot, Beta
        read(20, *) NlxTherm, NlyTherm, LtrotTherm
        read(20, *) Nwrap, Nbin, Nsweep
        read(20, *) is_tau, Nthermal
        read(20, *) is_warm, Nwarm
        read(20, *) is_global, Nglobal
        read(20, *) absolute
        close(20)
        Lq = Nlx * Nly
        Ndim = Lq
        Dtau = Beta / dble(Ltrot)
        if (mod(Ltrot, Nwrap) /= 0) then
            write(6, *) "Ltrot must be a multiple of Nwrap"; stop 1
        endif
        Nst = Ltrot / Nwrap
    end subroutine read_input

    ! Model_init (src/model.f90:16-36): one chain per line of seeds.txt
    subroutine Model_init(iseed)
        integer, intent(out) :: iseed
        type(z2_params) :: p
        type(z2_chain_desc), allocatable :: d(:)
        integer, allocatable :: seeds(:)
        integer :: k, ios, s, conf0
        call read_input()
        if (LtrotTherm > Ltrot) then
            write(6, *) "LtrotTherm must not exceed Ltrot (conf_transfer tiles the thermalised field)"; stop 1
        endif
        open(unit=30, file='confin.txt', status='old')
        read(30, *) conf0
        close(30)
        allocate(seeds(4096))
        n_chains = 0
        open(unit=10, file='seeds.txt', status='old')
        do
            read(10, *, iostat=ios) s
            if (ios /= 0) exit
            n_chains = n_chains + 1
            seeds(n_chains) = s
            if (n_chains == size(seeds) .or. n_chains == max_chains) exit
        enddo
        close(10)
        if (n_chains == 0) then
            write(6, *) "seeds.txt holds no seed"; stop 1
        endif
        allocate(d(n_chains))
        do k = 1, n_chains
            d(k) = z2_chain_desc(RT, J, h, mu, Q, Beta, int(seeds(k), c_int32_t), int(absolute, c_int32_t))
        enddo
        p = z2_params(int(Nlx, c_int32_t), int(Nly, c_int32_t), int(Ltrot, c_int32_t), &
                      int(Nwrap, c_int32_t), int(Nsweep, c_int32_t), int(precision_mode, c_int32_t))
        call z2_check(z2_ctx_create(p, int(n_chains, c_int32_t), d, ctx), "z2_ctx_create")
        ctx_params = p
        ctx_chains = d
        if (conf0 == 0) then
            if (NlxTherm == Nlx .and. NlyTherm == Nly .and. LtrotTherm == Ltrot) then
                call z2_check(z2_init_fields(ctx), "z2_init_fields")
            else  ! conf_set at the thermalisation size, then conf_transfer (src/fields.f90:429-507)
                call z2_check(z2_init_fields_therm(ctx, int(NlxTherm, c_int32_t), int(NlyTherm, c_int32_t), &
                              int(LtrotTherm, c_int32_t)), "z2_init_fields_therm")
            endif
        else
            call conf_restart()
        endif
        call z2_check(z2_get_seed(ctx, 0_c_int32_t, iseed), "z2_get_seed")
    end subroutine Model_init

    ! restart from confin.txt (conf_therm_in, src/fields.f90:227-309): per chain its
    ! seed, sigma(ii, nt) with nt outer, then lambda(1:Lq), at the thermalisation
    ! size; a missing lambda is drawn from the chain's seed (lambda_random_fill,
    ! src/fields.f90:559-603); conf_transfer tiles them onto the run's lattice
    subroutine conf_restart()
        real(c_double), allocatable :: sig(:), lam(:)
        integer :: k, i, ios, LqT
        integer(c_int32_t) :: s
        LqT = NlxTherm * NlyTherm
        allocate(sig(2 * LqT * LtrotTherm), lam(LqT))
        open(unit=30, file='confin.txt', status='old')
        do k = 0, n_chains - 1
            read(30, *) s
            do i = 1, size(sig)
                read(30, *) sig(i)
            enddo
            ios = 0
            do i = 1, LqT
                read(30, *, iostat=ios) lam(i)
                if (ios /= 0) exit
            enddo
            if (ios /= 0) call lambda_fill(lam, s)
            call z2_check(z2_set_sigma_therm(ctx, int(k, c_int32_t), int(NlxTherm, c_int32_t), &
                          int(NlyTherm, c_int32_t), int(LtrotTherm, c_int32_t), sig, lam, s), "conf_transfer")
        enddo
        close(30)
    end subroutine conf_restart

    ! lambda_random_fill + lambda_enforce_sector with the ranf stream of the seed
    subroutine lambda_fill(lam, iq)
        real(c_double), intent(out) :: lam(:)
        integer(c_int32_t), intent(inout) :: iq
        integer(kind=8) :: x
        real(kind=8) :: prod
        integer :: i
        prod = 1.d0
        x = iand(int(iq, 8), 2147483647_8)
        do i = 1, size(lam)
            x = mod(x * 48828125_8, 2147483648_8)
            lam(i) = -1.d0
            if (dble(x) / 2147483648.d0 >= 0.5d0) lam(i) = 1.d0
            prod = prod * lam(i)
        enddo
        if ((Q >= 0.d0 .and. prod < 0.d0) .or. (Q < 0.d0 .and. prod > 0.d0)) lam(size(lam)) = -lam(size(lam))
        iq = int(x, c_int32_t)
    end subroutine lambda_fill

    ! Model_clear -> conf_out (src/model.f90:38-45, src/fields.f90:509-557):
    ! seed, sigma(ii, nt), lambda(i) of every chain
    subroutine Model_clear(iseed)
        integer, intent(in) :: iseed
        real(c_double), allocatable :: sig(:), lam(:)
        integer(c_int32_t) :: s
        integer :: k, i
        allocate(sig(2 * Lq * Ltrot), lam(Lq))
        open(unit=35, file='confout.txt', status='unknown')
        do k = 0, n_chains - 1
            call z2_check(z2_get_seed(ctx, int(k, c_int32_t), s), "z2_get_seed")
            call z2_check(z2_get_sigma(ctx, int(k, c_int32_t), sig), "z2_get_sigma")
            call z2_check(z2_get_lambda(ctx, int(k, c_int32_t), lam), "z2_get_lambda")
            write(35, *) s
            do i = 1, size(sig)
                write(35, *) sig(i)
            enddo
            do i = 1, size(lam)
                write(35, *) lam(i)
            enddo
        enddo
        close(35)
        call z2_check(z2_ctx_destroy(ctx), "z2_ctx_destroy")
        ctx = c_null_ptr
        if (iseed == 0) continue
    end subroutine Model_clear
end module DQMC_Model_mod

module ProcessMatrix
    implicit none
    ! device-resident (src/process_matrix.f90:6-44): the Fortran objects carry the
    ! drift figures and whether they stand for the spin-down copy
    type :: Propagator
        integer :: chain = 0
        logical :: down = .false.
        real(kind=8) :: Xmaxm = 0.d0, Xmeanm = 0.d0
    contains
        procedure :: make => Prop_make
    end type Propagator
    type :: PropGreen
        logical :: down = .false.
        real(kind=8) :: Xmaxm(3) = 0.d0, Xmeanm(3) = 0.d0
    contains
        procedure :: make => Propgr_make
    end type PropGreen
    type :: WrapList
        integer :: chain = 0
        logical :: down = .false.
    contains
        procedure :: make => Wrlist_make
    end type WrapList
    integer, save, private :: n_prop = 0, n_propgr = 0, n_wrlist = 0
contains
    subroutine Prop_make(this)
        class(Propagator), intent(inout) :: this
        n_prop = n_prop + 1
        this%down = mod(n_prop, 2) == 0
        this%Xmaxm = 0.d0
        this%Xmeanm = 0.d0
    end subroutine Prop_make
    subroutine Propgr_make(this)
        class(PropGreen), intent(inout) :: this
        n_propgr = n_propgr + 1
        this%down = mod(n_propgr, 2) == 0
        this%Xmaxm = 0.d0
        this%Xmeanm = 0.d0
    end subroutine Propgr_make
    subroutine Wrlist_make(this)
        class(WrapList), intent(inout) :: this
        n_wrlist = n_wrlist + 1
        this%down = mod(n_wrlist, 2) == 0
        this%chain = 0
    end subroutine Wrlist_make
end module ProcessMatrix

module Stabilize_mod
    use iso_c_binding
    use z2dqmc_c
    use DQMC_Model_mod, only: ctx
    use ProcessMatrix
    implicit none
contains
    subroutine Stabilize_init()
    end subroutine Stabilize_init
    subroutine Stabilize_clear()
    end subroutine Stabilize_clear
    ! src/stabilization.f90:374-476 (all chains at once)
    subroutine Wrap_pre(Prop, WrList, nt)
        class(Propagator), intent(inout) :: Prop
        class(WrapList), intent(inout) :: WrList
        integer, intent(in) :: nt
        if (Prop%down) return
        call z2_check(z2_wrap_pre(ctx, int(nt, c_int32_t)), "Wrap_pre")
    end subroutine Wrap_pre
    subroutine Wrap_L(Prop, WrList, nt, flag)
        class(Propagator), intent(inout) :: Prop
        class(WrapList), intent(inout) :: WrList
        integer, intent(in) :: nt
        character(len=*), optional, intent(in) :: flag
        integer(c_int32_t) :: f
        if (Prop%down) return
        f = 0
        if (present(flag)) f = 1
        call z2_check(z2_wrap_L(ctx, int(nt, c_int32_t), f), "Wrap_L")
    end subroutine Wrap_L
    subroutine Wrap_R(Prop, WrList, nt, flag)
        class(Propagator), intent(inout) :: Prop
        class(WrapList), intent(inout) :: WrList
        integer, intent(in) :: nt
        character(len=*), optional, intent(in) :: flag
        integer(c_int32_t) :: f
        if (Prop%down) return
        f = 0
        if (present(flag)) f = 1
        call z2_check(z2_wrap_R(ctx, int(nt, c_int32_t), f), "Wrap_R")
    end subroutine Wrap_R
    ! src/stabilization.f90:478-514 on the time-displaced state
    subroutine Wrap_tau(Prop, PropGr, WrList, nt)
        class(Propagator), intent(inout) :: Prop
        class(PropGreen), intent(inout) :: PropGr
        class(WrapList), intent(in) :: WrList
        integer, intent(in) :: nt
        if (Prop%down) return
        call z2_check(z2_wrap_tau(ctx, int(nt, c_int32_t)), "Wrap_tau")
    end subroutine Wrap_tau
end module Stabilize_mod

module ObserEqual_mod
    use iso_c_binding
    use z2dqmc_c
    use DQMC_Model_mod, only: ctx, n_chains
    use ProcessMatrix
    implicit none
    ! src/obser_equal.f90:7-20 (chain-averaged, as after the per-bin MPI_REDUCE)
    type, public :: ObserEqual
        real(kind=8) :: flux_avg, chi_flux, pair_sc_q0, cdw_pi, cdw_shift, cdw_ratio, density
    contains
        procedure :: make => Obs_equal_make
        procedure :: reset => Obs_equal_reset
        procedure :: calc => Obs_equal_calc
        procedure :: ave => Obs_equal_ave
        procedure :: collect => Obs_equal_collect
    end type ObserEqual
contains
    subroutine Obs_equal_make(this)
        class(ObserEqual), intent(inout) :: this
        call obs_equal_zero(this)
    end subroutine Obs_equal_make
    subroutine obs_equal_zero(this)
        class(ObserEqual), intent(inout) :: this
        this%flux_avg = 0.d0; this%chi_flux = 0.d0; this%pair_sc_q0 = 0.d0
        this%cdw_pi = 0.d0; this%cdw_shift = 0.d0; this%cdw_ratio = 0.d0; this%density = 0.d0
    end subroutine obs_equal_zero
    ! src/obser_equal.f90:33-48
    subroutine Obs_equal_reset(this)
        class(ObserEqual), intent(inout) :: this
        call obs_equal_zero(this)
        if (c_associated(ctx)) call z2_check(z2_obs_reset(ctx), "Obs_equal%reset")
    end subroutine Obs_equal_reset
    ! src/obser_equal.f90:73-149, all chains at slice ntau
    subroutine Obs_equal_calc(this, PropU, PropD, ntau)
        class(ObserEqual), intent(inout) :: this
        class(Propagator), intent(in) :: PropU, PropD
        integer, intent(in) :: ntau
        call z2_check(z2_obs_equal_calc(ctx, int(ntau, c_int32_t)), "Obs_equal%calc")
    end subroutine Obs_equal_calc
    ! Obs_equal%ave (src/obser_equal.f90:50-71) on the engine, then the chain average
    subroutine Obs_equal_ave(this, Nobs)
        class(ObserEqual), intent(inout) :: this
        integer, intent(in) :: Nobs
        call z2_check(z2_obs_equal_ave(ctx, int(Nobs, c_int32_t)), "Obs_equal%ave")
        call this%collect()
    end subroutine Obs_equal_ave
    ! the per-bin rank average of m_process_obs_equal (src/fourier_trans.f90:280-315)
    ! over the chains of the per-chain averages the engine holds
    subroutine Obs_equal_collect(this)
        class(ObserEqual), intent(inout) :: this
        real(c_double) :: o(7)
        integer :: k
        call obs_equal_zero(this)
        do k = 0, n_chains - 1
            call z2_check(z2_get_obs_equal(ctx, int(k, c_int32_t), o), "z2_get_obs_equal")
            this%flux_avg = this%flux_avg + o(1)
            this%chi_flux = this%chi_flux + o(2)
            this%pair_sc_q0 = this%pair_sc_q0 + o(3)
            this%cdw_pi = this%cdw_pi + o(4)
            this%cdw_shift = this%cdw_shift + o(5)
            this%cdw_ratio = this%cdw_ratio + o(6)
            this%density = this%density + o(7)
        enddo
        this%flux_avg = this%flux_avg / dble(n_chains)
        this%chi_flux = this%chi_flux / dble(n_chains)
        this%pair_sc_q0 = this%pair_sc_q0 / dble(n_chains)
        this%cdw_pi = this%cdw_pi / dble(n_chains)
        this%cdw_shift = this%cdw_shift / dble(n_chains)
        this%cdw_ratio = this%cdw_ratio / dble(n_chains)
        this%density = this%density / dble(n_chains)
    end subroutine Obs_equal_collect
end module ObserEqual_mod

module ObserTau_mod
    use iso_c_binding
    use z2dqmc_c
    use DQMC_Model_mod, only: ctx, n_chains, Lq, Ltrot
    use ProcessMatrix
    implicit none
    ! src/obser_tau.f90:7-16: Lq x Ltrot correlators at displacement imj and slice
    type, public :: ObserTau
        complex(kind=8), allocatable :: spin_tau(:,:), pair_s_tau(:,:), pair_d_tau(:,:)
        complex(kind=8), allocatable :: charge_s_tau(:,:), charge_d_tau(:,:), single_tau(:,:), curxx_tau(:,:)
    contains
        procedure :: make => Obs_tau_make
        procedure :: reset => Obs_tau_reset
        procedure :: calc => Obs_tau_calc
        procedure :: ave => Obs_tau_ave
        procedure :: collect => Obs_tau_collect
    end type ObserTau
    ! the engine's accumulators exist (z2_enable_tau has run)
    logical :: tau_on_device = .false.
contains
    subroutine Obs_tau_make(this)
        class(ObserTau), intent(inout) :: this
        allocate(this%spin_tau(Lq, Ltrot), this%pair_s_tau(Lq, Ltrot), this%pair_d_tau(Lq, Ltrot))
        allocate(this%charge_s_tau(Lq, Ltrot), this%charge_d_tau(Lq, Ltrot), this%single_tau(Lq, Ltrot))
        allocate(this%curxx_tau(Lq, Ltrot))
        call obs_tau_zero(this)
    end subroutine Obs_tau_make
    subroutine obs_tau_zero(this)
        class(ObserTau), intent(inout) :: this
        this%spin_tau = (0.d0, 0.d0); this%pair_s_tau = (0.d0, 0.d0); this%pair_d_tau = (0.d0, 0.d0)
        this%charge_s_tau = (0.d0, 0.d0); this%charge_d_tau = (0.d0, 0.d0); this%single_tau = (0.d0, 0.d0)
        this%curxx_tau = (0.d0, 0.d0)
    end subroutine obs_tau_zero
    ! src/obser_tau.f90:26-35
    subroutine Obs_tau_reset(this)
        class(ObserTau), intent(inout) :: this
        call obs_tau_zero(this)
        if (tau_on_device) call z2_check(z2_obs_tau_reset(ctx), "Obs_tau%reset")
    end subroutine Obs_tau_reset
    ! src/obser_tau.f90:56-133, all chains at slice ntau
    subroutine Obs_tau_calc(this, PropGrU, PropGrD, ntau)
        class(ObserTau), intent(inout) :: this
        class(PropGreen), intent(in) :: PropGrU, PropGrD
        integer, intent(in) :: ntau
        call z2_check(z2_obs_tau_calc(ctx, int(ntau, c_int32_t)), "Obs_tau%calc")
    end subroutine Obs_tau_calc
    ! Obs_tau%ave (src/obser_tau.f90:37-47) on the engine, then the chain average
    subroutine Obs_tau_ave(this, Nobst)
        class(ObserTau), intent(inout) :: this
        integer, intent(in) :: Nobst
        call z2_check(z2_obs_tau_ave(ctx, int(Nobst, c_int32_t)), "Obs_tau%ave")
        call this%collect()
    end subroutine Obs_tau_ave
    ! the rank average of m_process_obs_tau (src/fourier_trans.f90:451-488) over the
    ! chains of the per-chain averages the engine holds
    subroutine Obs_tau_collect(this)
        class(ObserTau), intent(inout) :: this
        real(c_double), allocatable :: o(:, :, :)
        integer :: k
        allocate(o(Lq, Ltrot, 7))
        call obs_tau_zero(this)
        do k = 0, n_chains - 1
            call z2_check(z2_get_obs_tau(ctx, int(k, c_int32_t), o), "z2_get_obs_tau")
            this%spin_tau = this%spin_tau + o(:, :, 1)
            this%pair_s_tau = this%pair_s_tau + o(:, :, 2)
            this%pair_d_tau = this%pair_d_tau + o(:, :, 3)
            this%charge_s_tau = this%charge_s_tau + o(:, :, 4)
            this%charge_d_tau = this%charge_d_tau + o(:, :, 5)
            this%single_tau = this%single_tau + o(:, :, 6)
            this%curxx_tau = this%curxx_tau + o(:, :, 7)
        enddo
        this%spin_tau = this%spin_tau / dcmplx(dble(n_chains), 0.d0)
        this%pair_s_tau = this%pair_s_tau / dcmplx(dble(n_chains), 0.d0)
        this%pair_d_tau = this%pair_d_tau / dcmplx(dble(n_chains), 0.d0)
        this%charge_s_tau = this%charge_s_tau / dcmplx(dble(n_chains), 0.d0)
        this%charge_d_tau = this%charge_d_tau / dcmplx(dble(n_chains), 0.d0)
        this%single_tau = this%single_tau / dcmplx(dble(n_chains), 0.d0)
        this%curxx_tau = this%curxx_tau / dcmplx(dble(n_chains), 0.d0)
    end subroutine Obs_tau_collect
end module ObserTau_mod

! Output of the time-displaced correlators (m_write_obs_tau, src/fourier_trans.f90:316-449):
! naive Fourier transform of the displacement index to the Lq lattice momenta
! (Fourier_R_to_K, src/lattice.f90:266-290), time integration to susceptibilities,
! files appended per bin.  Momentum and displacement tables as src/lattice.f90:157-175.
module FourierTrans_mod
    use DQMC_Model_mod, only: Nlx, Nly, Lq, Ltrot, Dtau, Beta
    use ObserTau_mod
    implicit none
    real(kind=8), allocatable, private :: xk_v(:,:), aimj_v(:,:), k_dot_r(:,:)
    integer, allocatable, private :: inv_n_list(:,:)
    type, public :: FourierTrans
    contains
        procedure, nopass :: init => Fourier_init
        procedure, nopass :: prtau => Fourier_prtau
    end type FourierTrans
contains
    subroutine Fourier_init()
        real(kind=8) :: pi
        integer :: ix, iy, ii, nk
        pi = acos(-1.d0)
        if (allocated(xk_v)) return
        allocate(xk_v(Lq, 2), aimj_v(Lq, 2), k_dot_r(Lq, Lq), inv_n_list(Nlx, Nly))
        do iy = 1, Nly
            do ix = 1, Nlx
                ii = (iy - 1) * Nlx + ix
                inv_n_list(ix, iy) = ii
                aimj_v(ii, 1) = dble(ix)
                aimj_v(ii, 2) = dble(iy)
                xk_v(ii, 1) = dble(ix - 1) * (2.d0 * pi) / dble(Nlx)
                xk_v(ii, 2) = dble(iy - 1) * (2.d0 * pi) / dble(Nly)
            enddo
        enddo
        do ii = 1, Lq
            do nk = 1, Lq
                k_dot_r(nk, ii) = xk_v(nk, 1) * aimj_v(ii, 1) + xk_v(nk, 2) * aimj_v(ii, 2)
            enddo
        enddo
    end subroutine Fourier_init

    ! gk(k, t) = sum over imj of exp(i k.r_imj) gr(imj, t) / Lq, imj outer
    subroutine r_to_k(gr, gk)
        complex(kind=8), intent(in) :: gr(:,:)
        complex(kind=8), intent(out) :: gk(:,:)
        integer :: nf, imj, nk
        gk = (0.d0, 0.d0)
        do nf = 1, size(gr, 2)
            do imj = 1, Lq
                do nk = 1, Lq
                    gk(nk, nf) = gk(nk, nf) + exp(dcmplx(0.d0, k_dot_r(nk, imj))) * gr(imj, nf)
                enddo
            enddo
        enddo
        gk = gk / dble(Lq)
    end subroutine r_to_k

    ! Dtau * sum over the slices (m_integrate_susc_2_mom)
    subroutine integrate_susc(gr, gk)
        complex(kind=8), intent(in) :: gr(:,:)
        complex(kind=8), intent(out) :: gk(:)
        integer :: nt, nk
        gk = (0.d0, 0.d0)
        do nt = 1, Ltrot
            do nk = 1, Lq
                gk(nk) = gk(nk) + gr(nk, nt)
            enddo
        enddo
        gk = gk * dcmplx(Dtau, 0.d0)
    end subroutine integrate_susc

    ! Dtau * sum over slices of exp(2 pi i (nt-1)(nw-1)/Ltrot) gr(1, nt) (m_integrate_susc_2_freq)
    subroutine integrate_susc_freq(gr, gw)
        complex(kind=8), intent(in) :: gr(:,:)
        complex(kind=8), intent(out) :: gw(:)
        integer :: nt, nw
        real(kind=8) :: pi
        pi = acos(-1.d0)
        gw = (0.d0, 0.d0)
        do nw = 1, Ltrot
            do nt = 1, Ltrot
                gw(nw) = gw(nw) + exp(dcmplx(0.d0, 2.d0 * pi * dble((nt - 1) * (nw - 1)) / dble(Ltrot))) * gr(1, nt)
            enddo
        enddo
        gw = gw * dcmplx(Dtau, 0.d0)
    end subroutine integrate_susc_freq

    subroutine write_reciprocal(gk, filek)
        complex(kind=8), intent(in) :: gk(:)
        character(len=*), intent(in) :: filek
        integer :: nk
        open(unit=20, file=filek, status='unknown', action='write', position='append')
        do nk = 1, Lq
            write(20, *) xk_v(nk, 1), xk_v(nk, 2)
            write(20, *) gk(nk)
        enddo
        close(20)
    end subroutine write_reciprocal

    subroutine write_k(gk, filek, idx)
        complex(kind=8), intent(in) :: gk(:)
        character(len=*), intent(in) :: filek
        integer, intent(in) :: idx
        open(unit=30, file=filek, status='unknown', action='write', position='append')
        write(30, *) real(gk(idx)), aimag(gk(idx))
        close(30)
    end subroutine write_k

    subroutine write_k_tau(gk, filek, idx)
        complex(kind=8), intent(in) :: gk(:,:)
        character(len=*), intent(in) :: filek
        integer, intent(in) :: idx
        integer :: nt
        open(unit=20, file=filek, status='unknown', action='write', position='append')
        write(20, *) xk_v(idx, 1), xk_v(idx, 2)
        do nt = 1, Ltrot
            write(20, *) real(gk(idx, nt)), aimag(gk(idx, nt))
        enddo
        close(20)
    end subroutine write_k_tau

    subroutine write_r_tau(gr, filek, idx)
        complex(kind=8), intent(in) :: gr(:,:)
        character(len=*), intent(in) :: filek
        integer, intent(in) :: idx
        integer :: nt
        complex(kind=8) :: tmp
        open(unit=20, file=filek, status='unknown', action='write', position='append')
        write(20, *) aimj_v(idx, 1), aimj_v(idx, 2)
        do nt = 1, Ltrot
            tmp = gr(idx, nt) / dble(Lq)
            write(20, *) real(tmp), aimag(tmp)
        enddo
        close(20)
    end subroutine write_r_tau

    subroutine write_w(gw, filek)
        complex(kind=8), intent(in) :: gw(:)
        character(len=*), intent(in) :: filek
        integer :: nw
        real(kind=8) :: pi
        pi = acos(-1.d0)
        open(unit=20, file=filek, status='unknown', action='write', position='append')
        do nw = 1, Ltrot
            write(20, *) 2.d0 * pi * dble(nw - 1) / Beta
            write(20, *) real(gw(nw)), aimag(gw(nw))
        enddo
        close(20)
    end subroutine write_w

    ! the files of m_write_obs_tau, in its order
    subroutine Fourier_prtau(Obs)
        class(ObserTau), intent(in) :: Obs
        complex(kind=8), allocatable :: gk2(:,:), gk1(:), gw(:)
        integer :: izero, ipi, iqx, iqy
        logical :: has_pi, has_qx, has_qy
        call Fourier_init()
        allocate(gk2(Lq, Ltrot), gk1(Lq), gw(Ltrot))
        izero = 1
        has_pi = mod(Nlx, 2) == 0 .and. mod(Nly, 2) == 0
        ipi = -1
        if (has_pi) ipi = inv_n_list(Nlx / 2 + 1, Nly / 2 + 1)
        has_qx = Nlx > 1
        has_qy = Nly > 1
        iqx = izero
        iqy = izero
        if (has_qx) iqx = inv_n_list(2, 1)
        if (has_qy) iqy = inv_n_list(1, 2)

        call r_to_k(Obs%spin_tau, gk2)
        call integrate_susc(gk2, gk1)
        call write_reciprocal(gk1, 'sdw_susc')
        if (has_pi) call write_k(gk1, 'sdw_pipi_susc', ipi)
        call write_k(gk1, 'sdw_zero_susc', izero)

        call r_to_k(Obs%charge_d_tau, gk2)
        if (has_pi) call write_k_tau(gk2, 'cdw_d_pipi_tau', ipi)
        call integrate_susc(gk2, gk1)
        call write_reciprocal(gk1, 'cdw_d_susc')
        if (has_pi) call write_k(gk1, 'cdw_d_pipi_susc', ipi)
        call write_k(gk1, 'cdw_d_zero_susc', izero)

        call r_to_k(Obs%charge_s_tau, gk2)
        if (has_pi) call write_k_tau(gk2, 'cdw_s_pipi_tau', ipi)
        call integrate_susc(gk2, gk1)
        call write_reciprocal(gk1, 'cdw_s_susc')
        if (has_pi) call write_k(gk1, 'cdw_s_pipi_susc', ipi)
        call write_k(gk1, 'cdw_s_zero_susc', izero)

        call write_r_tau(Obs%single_tau, 'g_r0_tau', Lq)
        call r_to_k(Obs%single_tau, gk2)
        call integrate_susc(gk2, gk1)
        call write_reciprocal(gk1, 'g_susc')
        if (has_pi) call write_k(gk1, 'g_pipi_susc', ipi)
        call write_k(gk1, 'g_zero_susc', izero)

        call r_to_k(Obs%curxx_tau, gk2)
        if (has_qx) call write_k_tau(gk2, 'curxx_L_tau', iqx)
        if (has_qy) call write_k_tau(gk2, 'curxx_T_tau', iqy)
        call write_k_tau(gk2, 'curxx_zero_tau', izero)
        call integrate_susc(gk2, gk1)
        call integrate_susc_freq(gk2, gw)
        call write_w(gw, 'curxx_freq')
        call write_reciprocal(gk1, 'curxx_susc')
        if (has_qx) call write_k(gk1, 'curxx_L_susc', iqx)
        if (has_qy) call write_k(gk1, 'curxx_T_susc', iqy)
        call write_k(gk1, 'curxx_zero_susc', izero)

        call r_to_k(Obs%pair_s_tau, gk2)
        call write_k_tau(gk2, 'pair_s_zero_tau', izero)
        call integrate_susc(gk2, gk1)
        call write_reciprocal(gk1, 'pair_s_susc')
        call write_k(gk1, 'pair_s_zero_susc', izero)

        call r_to_k(Obs%pair_d_tau, gk2)
        call write_k_tau(gk2, 'pair_d_zero_tau', izero)
        call integrate_susc(gk2, gk1)
        call write_reciprocal(gk1, 'pair_d_susc')
        call write_k(gk1, 'pair_d_zero_susc', izero)
    end subroutine Fourier_prtau
end module FourierTrans_mod

! src/multiply.f90:8-130: which matrices get B or B^{-1} from which side at each
! slice; all chains at slice nt (sigma and lambda live on the GPU: the array
! arguments are kept for the reference's signatures)
module Multiply_mod
    use iso_c_binding
    use z2dqmc_c
    use DQMC_Model_mod, only: ctx, Lq, Ltrot
    use ProcessMatrix
    implicit none
contains
    subroutine propK_pre(PropU, PropD, sigma, nt)
        class(Propagator), intent(inout) :: PropU, PropD
        real(kind=8), dimension(2*Lq, Ltrot), intent(in) :: sigma
        integer, intent(in) :: nt
        call z2_check(z2_prop_K_pre(ctx, int(nt, c_int32_t)), "propK_pre")
    end subroutine propK_pre
    subroutine propT_pre(PropU, PropD, lambda, nt)
        class(Propagator), intent(inout) :: PropU, PropD
        real(kind=8), dimension(Lq), intent(in) :: lambda
        integer, intent(in) :: nt
        call z2_check(z2_prop_T_pre(ctx, int(nt, c_int32_t)), "propT_pre")
    end subroutine propT_pre
    subroutine propT_L(PropU, PropD, lambda, nt)
        class(Propagator), intent(inout) :: PropU, PropD
        real(kind=8), dimension(Lq), intent(in) :: lambda
        integer, intent(in) :: nt
        call z2_check(z2_prop_T_L(ctx, int(nt, c_int32_t)), "propT_L")
    end subroutine propT_L
    subroutine propT_R(PropU, PropD, lambda, nt)
        class(Propagator), intent(inout) :: PropU, PropD
        real(kind=8), dimension(Lq), intent(in) :: lambda
        integer, intent(in) :: nt
        call z2_check(z2_prop_T_R(ctx, int(nt, c_int32_t)), "propT_R")
    end subroutine propT_R
    subroutine propgrK_R(PropU, PropD, PropgrU, PropgrD, sigma, nt)
        class(Propagator), intent(inout) :: PropU, PropD
        class(PropGreen), intent(inout) :: PropgrU, PropgrD
        real(kind=8), dimension(2*Lq, Ltrot), intent(in) :: sigma
        integer, intent(in) :: nt
        call z2_check(z2_propgr_K_R(ctx, int(nt, c_int32_t)), "propgrK_R")
    end subroutine propgrK_R
    subroutine propgrT_R(PropU, PropD, PropgrU, PropgrD, lambda, nt)
        class(Propagator), intent(inout) :: PropU, PropD
        class(PropGreen), intent(inout) :: PropgrU, PropgrD
        real(kind=8), dimension(Lq), intent(in) :: lambda
        integer, intent(in) :: nt
        call z2_check(z2_propgr_T_R(ctx, int(nt, c_int32_t)), "propgrT_R")
    end subroutine propgrT_R
end module Multiply_mod

! src/localK.f90:20-43, 347-398: the link-flip sweep of one slice and its wraps,
! all chains; each chain's ranf seed lives on the GPU (LocalK_prop_L/R leave the
! iseed argument as it is; Local_sweep returns chain 0's)
module LocalK_mod
    use iso_c_binding
    use z2dqmc_c
    use DQMC_Model_mod, only: ctx
    use ProcessMatrix
    implicit none
contains
    subroutine LocalK_init()
    end subroutine LocalK_init
    subroutine LocalK_clear()
    end subroutine LocalK_clear
    ! Acc_Kl%reset: the engine's acceptance counters are cumulative (z2_get_counters)
    subroutine LocalK_reset()
    end subroutine LocalK_reset
    subroutine LocalK_prop_L(PropU, PropD, iseed, nt)
        class(Propagator), intent(inout) :: PropU, PropD
        integer, intent(inout) :: iseed
        integer, intent(in) :: nt
        call z2_check(z2_localK_prop_L(ctx, int(nt, c_int32_t)), "LocalK_prop_L")
    end subroutine LocalK_prop_L
    subroutine LocalK_prop_R(PropU, PropD, iseed, nt)
        class(Propagator), intent(inout) :: PropU, PropD
        integer, intent(inout) :: iseed
        integer, intent(in) :: nt
        call z2_check(z2_localK_prop_R(ctx, int(nt, c_int32_t)), "LocalK_prop_R")
    end subroutine LocalK_prop_R
end module LocalK_mod

! src/dynamics.f90:1-95: the time-displaced Green's functions propagated over all
! slices from the start of a right pass, re-stabilised by Wrap_tau
module Dynamics_mod
    use iso_c_binding
    use z2dqmc_c
    use DQMC_Model_mod, only: ctx, n_chains, Ltrot, Nwrap, RT, Lq
    use ProcessMatrix
    use Stabilize_mod
    use Multiply_mod
    use ObserTau_mod
    implicit none
    type :: Dynamics
    contains
        procedure, nopass :: init => Dyn_init
        procedure, nopass :: clear => Dyn_clear
        procedure, nopass :: reset => Dyn_reset
        procedure, nopass :: sweep_R => Dyn_sweep_R
    end type Dynamics
    type(Propagator), allocatable :: Prop_dU, Prop_dD
    type(PropGreen), allocatable :: PropGrU, PropGrD
    real(kind=8), allocatable :: sig_arg(:, :), lam_arg(:)
    private :: Prop_dU, Prop_dD, PropGrU, PropGrD, sig_arg, lam_arg
contains
    subroutine Dyn_init()
        allocate(Prop_dU); call Prop_dU%make()
        allocate(Prop_dD); call Prop_dD%make()
        allocate(PropGrU); call PropGrU%make()
        allocate(PropGrD); call PropGrD%make()
        allocate(sig_arg(2 * Lq, Ltrot), lam_arg(Lq))
        sig_arg = 0.d0; lam_arg = 0.d0
    end subroutine Dyn_init
    subroutine Dyn_clear()
        deallocate(Prop_dU, Prop_dD, PropGrU, PropGrD, sig_arg, lam_arg)
    end subroutine Dyn_clear
    ! src/dynamics.f90:41-48
    subroutine Dyn_reset(PropU, PropD)
        class(Propagator), intent(in) :: PropU, PropD
        call z2_check(z2_dyn_reset(ctx), "Dyn%reset")
    end subroutine Dyn_reset
    ! src/dynamics.f90:50-62
    subroutine Dyn_sweep_R(Obs, WrU, WrD)
        class(ObserTau), intent(inout) :: Obs
        class(WrapList), intent(in) :: WrU, WrD
        integer :: nt
        do nt = 1, Ltrot
            call Obs%calc(PropGrU, PropGrD, nt)
            if (RT > 0.d0) call propgrK_R(Prop_dU, Prop_dD, PropGrU, PropGrD, sig_arg, nt)
            call propgrT_R(Prop_dU, Prop_dD, PropGrU, PropGrD, lam_arg, nt)
            if (mod(nt, Nwrap) == 0) call Wrap_tau(Prop_dU, PropGrU, WrU, nt)
            if (mod(nt, Nwrap) == 0) call Wrap_tau(Prop_dD, PropGrD, WrD, nt)
        enddo
    end subroutine Dyn_sweep_R
end module Dynamics_mod

module LocalSweep_mod
    use iso_c_binding
    use z2dqmc_c
    use DQMC_Model_mod
    use ProcessMatrix
    use ObserEqual_mod
    use ObserTau_mod
    use Stabilize_mod
    use Multiply_mod
    use LocalK_mod
    use Dynamics_mod
    implicit none
    type :: LocalSweep
    contains
        procedure :: init => Local_sweep_init
        procedure :: clear => Local_sweep_clear
        procedure :: pre => Local_sweep_pre
        procedure :: sweep => Local_sweep
        procedure :: therm => Local_sweep_therm
        procedure, nopass :: ctrl_print_l => Local_control_print
        procedure, nopass :: ctrl_print_t => Therm_control_print
    end type LocalSweep
    type(ObserEqual), allocatable :: Obs_equal
    type(ObserTau), allocatable :: Obs_tau
    type(Dynamics) :: Dyn
    ! .false.: one batched engine call per pre / sweep (z2_pre, z2_sweep_tau);
    ! .true.: the reference's procedure sequence (src/local_sweep.f90:87-195), one
    ! engine call per Multiply_mod / LocalK_mod / Stabilize_mod / Obs procedure
    logical :: fine_grained = .false.
    real(kind=8), allocatable :: sig_arg(:, :), lam_arg(:)
contains
    subroutine Local_sweep_init(this)
        class(LocalSweep), intent(inout) :: this
        allocate(Obs_equal)
        call Obs_equal%make()
        if (is_tau) then
            allocate(Obs_tau)
            call Obs_tau%make()
            call z2_check(z2_enable_tau(ctx), "z2_enable_tau")
            tau_on_device = .true.
            call Dyn%init()
        endif
        allocate(sig_arg(2 * Lq, Ltrot), lam_arg(Lq))
        sig_arg = 0.d0; lam_arg = 0.d0
    end subroutine Local_sweep_init
    subroutine Local_sweep_clear(this)
        class(LocalSweep), intent(inout) :: this
        deallocate(Obs_equal)
        if (allocated(Obs_tau)) then
            deallocate(Obs_tau)
            call Dyn%clear()
        endif
        deallocate(sig_arg, lam_arg)
    end subroutine Local_sweep_clear
    ! src/local_sweep.f90:87-104
    subroutine Local_sweep_pre(this, PropU, PropD, WrU, WrD)
        class(LocalSweep), intent(inout) :: this
        class(Propagator), intent(inout) :: PropU, PropD
        class(WrapList), intent(inout) :: WrU, WrD
        integer :: nt
        if (.not. fine_grained) then
            call z2_check(z2_pre(ctx), "z2_pre")
            return
        endif
        call z2_check(z2_sweep_reset(ctx, 0_c_int32_t), "Local_sweep_reset")
        call Wrap_pre(PropU, WrU, 0)
        call Wrap_pre(PropD, WrD, 0)
        do nt = 1, Ltrot
            if (RT > 0.d0) call propK_pre(PropU, PropD, sig_arg, nt)
            call propT_pre(PropU, PropD, lam_arg, nt)
            if (mod(nt, Nwrap) == 0) then
                call Wrap_pre(PropU, WrU, nt)
                call Wrap_pre(PropD, WrD, nt)
            endif
        enddo
    end subroutine Local_sweep_pre
    ! Local_sweep_therm (src/local_sweep.f90:73-85), all chains
    subroutine Local_sweep_therm(this, iseed)
        class(LocalSweep), intent(inout) :: this
        integer, intent(inout) :: iseed
        integer(c_int32_t) :: s
        call z2_check(z2_therm(ctx), "z2_therm")
        call z2_check(z2_get_seed(ctx, 0_c_int32_t, s), "z2_get_seed")
        iseed = s
    end subroutine Local_sweep_therm
    ! src/local_sweep.f90:106-124
    subroutine Local_sweep_L(PropU, PropD, WrU, WrD, iseed, Nobs)
        class(Propagator), intent(inout) :: PropU, PropD
        class(WrapList), intent(inout) :: WrU, WrD
        integer, intent(inout) :: iseed, Nobs
        integer :: nt
        do nt = Ltrot, 1, -1
            if (mod(nt, Nwrap) == 0) then
                call Wrap_L(PropU, WrU, nt, "S")
                call Wrap_L(PropD, WrD, nt, "S")
            endif
            call Obs_equal%calc(PropU, PropD, nt)
            Nobs = Nobs + 1
            call propT_L(PropU, PropD, lam_arg, nt)
            call LocalK_prop_L(PropU, PropD, iseed, nt)
        enddo
        call Wrap_L(PropU, WrU, 0, "S")
        call Wrap_L(PropD, WrD, 0, "S")
    end subroutine Local_sweep_L
    ! src/local_sweep.f90:126-150
    subroutine Local_sweep_R(PropU, PropD, WrU, WrD, iseed, toggle, Nobs, Nobst)
        class(Propagator), intent(inout) :: PropU, PropD
        class(WrapList), intent(inout) :: WrU, WrD
        logical, intent(in) :: toggle
        integer, intent(inout) :: iseed, Nobs, Nobst
        integer :: nt
        if (toggle) then
            call Dyn%reset(PropU, PropD)
            call Dyn%sweep_R(Obs_tau, WrU, WrD)
            Nobst = Nobst + 1
        endif
        call Wrap_R(PropU, WrU, 0, "S")
        call Wrap_R(PropD, WrD, 0, "S")
        do nt = 1, Ltrot
            call LocalK_prop_R(PropU, PropD, iseed, nt)
            call propT_R(PropU, PropD, lam_arg, nt)
            if (mod(nt, Nwrap) == 0) then
                call Wrap_R(PropU, WrU, nt, "S")
                call Wrap_R(PropD, WrD, nt, "S")
            endif
            call Obs_equal%calc(PropU, PropD, nt)
            Nobs = Nobs + 1
        enddo
    end subroutine Local_sweep_R
    ! src/local_sweep.f90:152-195 (all chains; batched: each chain in its own
    ! direction, the engine keeps every chain's is_beta; fine-grained: all chains in
    ! the direction is_beta; is_beta and iseed return chain 0's)
    subroutine Local_sweep(this, PropU, PropD, WrU, WrD, iseed, is_beta, toggle)
        class(LocalSweep), intent(inout) :: this
        class(Propagator), intent(inout) :: PropU, PropD
        class(WrapList), intent(inout) :: WrU, WrD
        integer, intent(inout) :: iseed
        logical, intent(inout) :: is_beta
        logical, intent(in) :: toggle
        integer(c_int32_t) :: b, s
        integer :: Nobs, Nobst, nsw, k
        real(c_double) :: xmax, xmean
        if (fine_grained) then
            do k = 1, n_chains - 1
                call z2_check(z2_get_is_beta(ctx, int(k, c_int32_t), b), "z2_get_is_beta")
                if ((b /= 0) .neqv. is_beta) then
                    write(6, *) "fine-grained sweep: chains in different directions (use the batched sweep)"
                    stop 1
                endif
            enddo
            call z2_check(z2_sweep_reset(ctx, merge(1_c_int32_t, 0_c_int32_t, toggle)), "Local_sweep_reset")
            Nobs = 0; Nobst = 0
            do nsw = 1, Nsweep
                if (is_beta) then
                    call Local_sweep_L(PropU, PropD, WrU, WrD, iseed, Nobs)
                    call Local_sweep_R(PropU, PropD, WrU, WrD, iseed, toggle, Nobs, Nobst)
                else
                    call Local_sweep_R(PropU, PropD, WrU, WrD, iseed, toggle, Nobs, Nobst)
                    call Local_sweep_L(PropU, PropD, WrU, WrD, iseed, Nobs)
                endif
            enddo
            call Obs_equal%ave(Nobs)
            if (toggle) call Obs_tau%ave(Nobst)
        else
            if (toggle) then
                ! with the time-displaced measurement (Dyn%sweep_R at every right pass)
                call z2_check(z2_sweep_tau(ctx, Z2_BETA_PER_CHAIN, 1_c_int32_t), "z2_sweep_tau")
                call Obs_tau%collect()
            else
                call z2_check(z2_sweep(ctx, Z2_BETA_PER_CHAIN), "z2_sweep")
            endif
            call Obs_equal%collect()
        endif
        call z2_check(z2_get_is_beta(ctx, 0_c_int32_t, b), "z2_get_is_beta")
        is_beta = b /= 0
        call z2_check(z2_get_seed(ctx, 0_c_int32_t, s), "z2_get_seed")
        iseed = s
        call z2_check(z2_get_drift(ctx, 0_c_int32_t, xmax, xmean), "z2_get_drift")
        PropU%Xmaxm = xmax; PropU%Xmeanm = xmean
        PropD%Xmaxm = xmax; PropD%Xmeanm = xmean
    end subroutine Local_sweep
    ! acceptance over all chains (src/local_sweep.f90:197-213)
    subroutine Local_control_print(toggle)
        logical, intent(in) :: toggle
        integer(c_int64_t) :: cnt(2), acc, prop
        integer :: k
        acc = 0; prop = 0
        do k = 0, n_chains - 1
            call z2_check(z2_get_counters(ctx, int(k, c_int32_t), cnt), "z2_get_counters")
            acc = acc + cnt(1); prop = prop + cnt(2)
        enddo
        write(50, *) 'Accept_Klocal_shift                            :', dble(acc) / dble(max(prop, 1_c_int64_t))
        write(50, *) 'Accept_lambda_local                            :', 0.d0
        if (toggle) continue
    end subroutine Local_control_print
    ! src/local_sweep.f90:215-227
    subroutine Therm_control_print()
        integer(c_int64_t) :: cnt(2), acc, prop
        integer :: k
        if (Nwarm <= 0) return
        acc = 0; prop = 0
        do k = 0, n_chains - 1
            call z2_check(z2_get_therm_counters(ctx, int(k, c_int32_t), cnt), "z2_get_therm_counters")
            acc = acc + cnt(1); prop = prop + cnt(2)
        enddo
        write(50, *) 'Thermalize Accept Ratio                        :', dble(acc) / dble(max(prop, 1_c_int64_t))
    end subroutine Therm_control_print
end module LocalSweep_mod

module GlobalUpdate_mod
    use iso_c_binding
    use z2dqmc_c
    use DQMC_Model_mod
    use ProcessMatrix
    implicit none
    ! src/global_update.f90:183-354 (Wolff cluster update of a bond's time line)
    type :: GlobalUpdate
    contains
        procedure :: init => Global_init
        procedure :: clear => Global_clear
        procedure :: sweep => Global_sweep
        procedure, nopass :: ctrl_print => Global_control_print
    end type GlobalUpdate
contains
    subroutine Global_init(this)
        class(GlobalUpdate), intent(inout) :: this
    end subroutine Global_init
    subroutine Global_clear(this)
        class(GlobalUpdate), intent(inout) :: this
    end subroutine Global_clear
    ! Global_sweep (src/global_update.f90:313-337): Nglobal trials per chain in the
    ! chain's own direction; is_beta and iseed return chain 0's
    subroutine Global_sweep(this, PropU, PropD, WrU, WrD, iseed, is_beta)
        class(GlobalUpdate), intent(inout) :: this
        class(Propagator), intent(inout) :: PropU, PropD
        class(WrapList), intent(inout) :: WrU, WrD
        integer, intent(inout) :: iseed
        logical, intent(inout) :: is_beta
        integer(c_int32_t) :: b, s
        call z2_check(z2_global_sweep(ctx, int(Nglobal, c_int32_t)), "z2_global_sweep")
        call z2_check(z2_get_is_beta(ctx, 0_c_int32_t, b), "z2_get_is_beta")
        is_beta = b /= 0
        call z2_check(z2_get_seed(ctx, 0_c_int32_t, s), "z2_get_seed")
        iseed = s
    end subroutine Global_sweep
    ! src/global_update.f90:339-354, averaged over the chains
    subroutine Global_control_print()
        integer(c_int64_t) :: cnt(3), acc, prop
        real(c_double) :: sz, szsum
        integer :: k
        acc = 0; prop = 0; szsum = 0.d0
        do k = 0, n_chains - 1
            call z2_check(z2_get_global_counters(ctx, int(k, c_int32_t), cnt, sz), "z2_get_global_counters")
            acc = acc + cnt(1); prop = prop + cnt(2); szsum = szsum + sz
        enddo
        write(50, *) 'Average size of Wolff cluster                  :', szsum / dble(n_chains * Nbin)
        write(50, *) 'Accept_Kglobal                                 :', dble(acc) / dble(max(prop, 1_c_int64_t))
    end subroutine Global_control_print
end module GlobalUpdate_mod
