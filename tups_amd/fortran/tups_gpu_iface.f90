!> ISO_C_BINDING interface to libtups_gpu (include/tups_gpu.h).
!>
!> Every C entry point is declared here with the reference's Fortran types:
!> complex(8) arrays pass as complex(c_double_complex) assumed-size dummies
!> (the reference's Fortran column-major layouts), integers are 32-bit like the
!> reference's default integer (makefile:17, MKL lp64).  tups_check mirrors the
!> reference's error convention: print the message and stop
!> (SRC/MPI_mod.f90:13-16, SRC/preprocessing.f90:95-98).
module tups_gpu_iface
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: TUPS_OK = 0, TUPS_EINVAL = -1, TUPS_EHIP = -2, &
                               TUPS_ENOMEM = -3, TUPS_ESTATE = -4, TUPS_ENONFINITE = -5, &
                               TUPS_ENOCONV = -6
  integer(c_int32_t), parameter :: TUPS_CH_D = 0, TUPS_CH_M = 1, TUPS_CH_S = 2, TUPS_CH_T = 3
  integer(c_int32_t), parameter :: TUPS_ARR_GAMMA = 0, TUPS_ARR_F = 1, TUPS_ARR_CHI = 2, TUPS_ARR_PHIR = 3, &
                                   TUPS_ARR_M = 4, TUPS_ARR_LAMBDA = 5
  integer(c_int32_t), parameter :: TUPS_BSE_LADDER_FP64 = 0, TUPS_BSE_LADDER_SPLIT = 1, &
                                   TUPS_BSE_CLOSED_LU = 2, TUPS_BSE_CLOSED_IR = 3
  integer(c_int32_t), parameter :: TUPS_K_BSE_LADDER = 0, TUPS_K_PHIR = 1, TUPS_K_BUILD_M = 2, &
                                   TUPS_K_PAE = 3, TUPS_K_GAMMA_UPDATE = 4, TUPS_K_PAE_GATHER = 5, &
                                   TUPS_K_SPLIT_PREP = 6, TUPS_K_SPLIT_GEMM = 7, TUPS_K_BSE_CLOSED = 8, &
                                   TUPS_K_CALC_CHI = 9, TUPS_K_BSE_BATCH = 10, TUPS_K_SDE = 11, &
                                   TUPS_K_GTAU = 12, TUPS_K_CONV = 13, TUPS_K_SUS = 14, TUPS_K_BSE_IR = 15

  type, bind(C) :: tups_dims
    integer(c_int32_t) :: nx, nf, nl, nr, nq, reserved
    real(c_double) :: beta, f_damping
  end type tups_dims

  type, bind(C) :: tups_conv_info
    real(c_double) :: sigma_avg, sigma_max
    integer(c_int32_t) :: kx, ky, nu, converged
    real(c_double) :: f_max(4), f_rel_err(4)
  end type tups_conv_info

  type, bind(C) :: tups_prms
    real(c_double) :: beta, u, n_particle
    integer(c_int32_t) :: nx, nf, neighbours, readinpick, ngrain, f_range, ite_max, dga
    integer(c_int32_t) :: use_sig_dmft, use_u2, save_f, oneshot, old, sde_old
    real(c_double) :: f_damping
    character(kind=c_char) :: pick_file(104), fname_l(104), fname_sig_dmft(104)
  end type tups_prms

  type, bind(C) :: tups_sizes
    integer(c_int32_t) :: nx, ny, nf, nl, nz, nr, nw, nibz, nq, nc, full_ibz, reserved
    real(c_double) :: alpha
  end type tups_sizes

  interface
    function tups_last_error() bind(C, name="tups_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function tups_last_error

    function tups_build_tables(dims, ff, ff_inv, signs, reps, exp_sqr, exp_qsr) &
        bind(C, name="tups_build_tables") result(rc)
      import :: tups_dims, c_double, c_int32_t, c_double_complex, c_int
      type(tups_dims), intent(in) :: dims
      real(c_double), intent(out) :: ff(*), ff_inv(*)
      integer(c_int32_t), intent(out) :: signs(*), reps(*)
      complex(c_double_complex), intent(out) :: exp_sqr(*), exp_qsr(*)
      integer(c_int) :: rc
    end function tups_build_tables

    function tups_create(dims, ctx) bind(C, name="tups_create") result(rc)
      import :: tups_dims, c_ptr, c_int
      type(tups_dims), intent(in) :: dims
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function tups_create

    function tups_destroy(ctx) bind(C, name="tups_destroy") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function tups_destroy

    function tups_query(ctx, sizes) bind(C, name="tups_query") result(rc)
      import :: c_ptr, c_int, tups_sizes
      type(c_ptr), value :: ctx
      type(tups_sizes), intent(out) :: sizes
      integer(c_int) :: rc
    end function tups_query

    function tups_get_tables(ctx, ff, ff_inv, signs, reps, exp_sqr, exp_qsr) &
        bind(C, name="tups_get_tables") result(rc)
      import :: c_ptr, c_double, c_int32_t, c_double_complex, c_int
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ff(*), ff_inv(*)
      integer(c_int32_t), intent(out) :: signs(*), reps(*)
      complex(c_double_complex), intent(out) :: exp_sqr(*), exp_qsr(*)
      integer(c_int) :: rc
    end function tups_get_tables

    function tups_set_tables(ctx, ff, ff_inv, signs, reps, exp_sqr, exp_qsr) &
        bind(C, name="tups_set_tables") result(rc)
      import :: c_ptr, c_double, c_int32_t, c_double_complex, c_int
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ff(*), ff_inv(*)
      integer(c_int32_t), intent(in) :: signs(*), reps(*)
      complex(c_double_complex), intent(in) :: exp_sqr(*), exp_qsr(*)
      integer(c_int) :: rc
    end function tups_set_tables

    function tups_get_q_map(ctx, qx, qy, qw) bind(C, name="tups_get_q_map") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: qx(*), qy(*), qw(*)
      integer(c_int) :: rc
    end function tups_get_q_map

    function tups_set_q_map(ctx, qx, qy, qw) bind(C, name="tups_set_q_map") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: qx(*), qy(*), qw(*)
      integer(c_int) :: rc
    end function tups_set_q_map

    function tups_upload(ctx, which, channel, host) bind(C, name="tups_upload") result(rc)
      import :: c_ptr, c_int32_t, c_double_complex, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which, channel
      complex(c_double_complex), intent(in) :: host(*)
      integer(c_int) :: rc
    end function tups_upload

    function tups_download(ctx, which, channel, host) bind(C, name="tups_download") result(rc)
      import :: c_ptr, c_int32_t, c_double_complex, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which, channel
      complex(c_double_complex), intent(out) :: host(*)
      integer(c_int) :: rc
    end function tups_download

    function tups_upload_chi_dense(ctx, pp, q0, nqs, dense) bind(C, name="tups_upload_chi_dense") result(rc)
      import :: c_ptr, c_int32_t, c_double_complex, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: pp, q0, nqs
      complex(c_double_complex), intent(in) :: dense(*)
      integer(c_int) :: rc
    end function tups_upload_chi_dense

    function tups_bse(ctx, mode) bind(C, name="tups_bse") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode
      integer(c_int) :: rc
    end function tups_bse

    function tups_set_refinement(ctx, tol, max_steps) bind(C, name="tups_set_refinement") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tol
      integer(c_int32_t), value :: max_steps
      integer(c_int) :: rc
    end function tups_set_refinement

    function tups_bse_solve_info(ctx, steps, residual) bind(C, name="tups_bse_solve_info") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: steps
      real(c_double), intent(out) :: residual
      integer(c_int) :: rc
    end function tups_bse_solve_info

    function tups_batch_create(nx, nq, nf, nl, beta, f_damping, batch) &
        bind(C, name="tups_batch_create") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double
      integer(c_int32_t), value :: nx, nq
      integer(c_int32_t), intent(in) :: nf(*), nl(*)
      real(c_double), value :: beta, f_damping
      type(c_ptr), intent(out) :: batch
      integer(c_int) :: rc
    end function tups_batch_create

    function tups_batch_destroy(batch) bind(C, name="tups_batch_destroy") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: batch
      integer(c_int) :: rc
    end function tups_batch_destroy

    function tups_batch_query(batch, vertex_elems, chi_elems, items) bind(C, name="tups_batch_query") result(rc)
      import :: c_ptr, c_int, c_int64_t, c_int32_t
      type(c_ptr), value :: batch
      integer(c_int64_t), intent(out) :: vertex_elems, chi_elems
      integer(c_int32_t), intent(out) :: items
      integer(c_int) :: rc
    end function tups_batch_query

    function tups_batch_upload(batch, which, channel, host) bind(C, name="tups_batch_upload") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double_complex
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: which, channel
      complex(c_double_complex), intent(in) :: host(*)
      integer(c_int) :: rc
    end function tups_batch_upload

    function tups_batch_download(batch, which, channel, host) bind(C, name="tups_batch_download") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double_complex
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: which, channel
      complex(c_double_complex), intent(out) :: host(*)
      integer(c_int) :: rc
    end function tups_batch_download

    function tups_batch_bse(batch, ms) bind(C, name="tups_batch_bse") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: batch
      type(c_ptr), value :: ms
      integer(c_int) :: rc
    end function tups_batch_bse

    function tups_batch_bse_split(batch, moduli, ms) bind(C, name="tups_batch_bse_split") result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: moduli
      type(c_ptr), value :: ms
      integer(c_int) :: rc
    end function tups_batch_bse_split

    function tups_batch_sync(batch) bind(C, name="tups_batch_sync") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: batch
      integer(c_int) :: rc
    end function tups_batch_sync

    function tups_batch_autotune(batch, variant, ms) bind(C, name="tups_batch_autotune") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_float
      type(c_ptr), value :: batch
      integer(c_int32_t), intent(out) :: variant
      real(c_float), intent(out) :: ms
      integer(c_int) :: rc
    end function tups_batch_autotune

    function tups_batch_set_variant(batch, variant) bind(C, name="tups_batch_set_variant") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: variant
      integer(c_int) :: rc
    end function tups_batch_set_variant

    ! uneven-box iteration: nf_w(Nw) boxes per bosonic frequency, nl_m(NIBZ) form factors per momentum
    function tups_set_product_forms(ctx, allow_three) bind(C, name="tups_set_product_forms") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: allow_three
      integer(c_int) :: rc
    end function tups_set_product_forms

    function tups_batch_set_product_forms(batch, allow_three) &
        bind(C, name="tups_batch_set_product_forms") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: allow_three
      integer(c_int) :: rc
    end function tups_batch_set_product_forms

    function tups_uneven_create(nx, nf_max, nf_w, nl_m, beta, f_damping, u) &
        bind(C, name="tups_uneven_create") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double
      integer(c_int32_t), value :: nx, nf_max
      integer(c_int32_t), intent(in) :: nf_w(*), nl_m(*)
      real(c_double), value :: beta, f_damping
      type(c_ptr), intent(out) :: u
      integer(c_int) :: rc
    end function tups_uneven_create

    function tups_uneven_destroy(u) bind(C, name="tups_uneven_destroy") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: u
      integer(c_int) :: rc
    end function tups_uneven_destroy

    function tups_uneven_parts(u, batch, ctx) bind(C, name="tups_uneven_parts") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: u
      type(c_ptr), intent(out) :: batch, ctx
      integer(c_int) :: rc
    end function tups_uneven_parts

    function tups_uneven_iteration(u, use_graph) bind(C, name="tups_uneven_iteration") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: u
      integer(c_int32_t), value :: use_graph
      integer(c_int) :: rc
    end function tups_uneven_iteration

    function tups_uneven_sync(u) bind(C, name="tups_uneven_sync") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: u
      integer(c_int) :: rc
    end function tups_uneven_sync

    function tups_sde(ctx, sigma, sigma_h, mu, ngrain, f_range, u, use_u2, sigma_out) &
        bind(C, name="tups_sde") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: sigma(*)
      type(c_ptr), value :: sigma_h
      real(c_double), value :: mu, u
      integer(c_int32_t), value :: ngrain, f_range, use_u2
      complex(c_double_complex), intent(out) :: sigma_out(*)
      integer(c_int) :: rc
    end function tups_sde

    function tups_set_sigma(ctx, sigma, sigma_h, mu, ngrain) bind(C, name="tups_set_sigma") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: sigma(*)
      type(c_ptr), value :: sigma_h
      real(c_double), value :: mu
      integer(c_int32_t), value :: ngrain
      integer(c_int) :: rc
    end function tups_set_sigma

    function tups_get_sigma(ctx, sigma) bind(C, name="tups_get_sigma") result(rc)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: sigma(*)
      integer(c_int) :: rc
    end function tups_get_sigma

    function tups_iteration_full(ctx, mode, use_graph, f_range, u, use_u2, n_particle) &
        bind(C, name="tups_iteration_full") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode, use_graph, f_range, use_u2
      real(c_double), value :: u, n_particle
      integer(c_int) :: rc
    end function tups_iteration_full

    ! calc_Gkw_Grt (SRC/parquet_chi.f90:18): mu in/out, Grt(Nx*Ngrain, Ny*Ngrain, Nf) out
    function tups_gkw_grt(ctx, u, n_particle, cal_chi, mu, steps, grt) bind(C, name="tups_gkw_grt") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), value :: u, n_particle
      integer(c_int32_t), value :: cal_chi
      real(c_double), intent(inout) :: mu
      integer(c_int32_t), intent(out) :: steps
      complex(c_double_complex), intent(out) :: grt(*)
      integer(c_int) :: rc
    end function tups_gkw_grt

    ! conv() (SRC/parquet_check.f90:39): pass c_null_ptr twice for the device Sigma pair
    function tups_conv(ctx, sigma_old, sigma, f_sigma, f_vert, info) bind(C, name="tups_conv") result(rc)
      import :: c_ptr, c_int, c_double, tups_conv_info
      type(c_ptr), value :: ctx, sigma_old, sigma
      real(c_double), value :: f_sigma, f_vert
      type(tups_conv_info), intent(out) :: info
      integer(c_int) :: rc
    end function tups_conv

    ! tups_conv with its partials in the library's fixed device array (no allocation)
    function tups_conv_noalloc(ctx, sigma_old, sigma, f_sigma, f_vert, info) &
        bind(C, name="tups_conv_noalloc") result(rc)
      import :: c_ptr, c_int, c_double, tups_conv_info
      type(c_ptr), value :: ctx, sigma_old, sigma
      real(c_double), value :: f_sigma, f_vert
      type(tups_conv_info), intent(out) :: info
      integer(c_int) :: rc
    end function tups_conv_noalloc

    ! calc_sus (SRC/parquet_sus.f90:18): sus(Nq, 4, 4) = sus_X_IBZ(q, ff) for X = d, m, s, t
    function tups_calc_sus(ctx, u, f_range, after_update, sus) bind(C, name="tups_calc_sus") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), value :: u
      integer(c_int32_t), value :: f_range, after_update
      complex(c_double_complex), intent(out) :: sus(*)
      integer(c_int) :: rc
    end function tups_calc_sus

    ! set_parameters (SRC/parquet_ini.f90:289): path is a NUL-terminated file name
    function tups_read_prms(path, prms) bind(C, name="tups_read_prms") result(rc)
      import :: c_char, c_int, tups_prms
      character(kind=c_char), intent(in) :: path(*)
      type(tups_prms), intent(out) :: prms
      integer(c_int) :: rc
    end function tups_read_prms

    function tups_prms_dims(prms, dims) bind(C, name="tups_prms_dims") result(rc)
      import :: c_int, tups_prms, tups_dims
      type(tups_prms), intent(in) :: prms
      type(tups_dims), intent(out) :: dims
      integer(c_int) :: rc
    end function tups_prms_dims

    ! calc_Gkw_Grt with cal_chi (SRC/parquet_chi.f90:216-430): Chi0_ph, Chi0_pp (Nred),
    ! Chi0_ph_FF, Chi0_pp_FF (Nred, 4)
    function tups_chi0(ctx, chi0_ph, chi0_pp, chi0_ph_ff, chi0_pp_ff) bind(C, name="tups_chi0") result(rc)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: chi0_ph(*), chi0_pp(*), chi0_ph_ff(*), chi0_pp_ff(*)
      integer(c_int) :: rc
    end function tups_chi0

    function tups_get_mu(ctx, mu, steps) bind(C, name="tups_get_mu") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: mu
      integer(c_int32_t), intent(out) :: steps
      integer(c_int) :: rc
    end function tups_get_mu

    function tups_iteration(ctx, mode, use_graph) bind(C, name="tups_iteration") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode, use_graph
      integer(c_int) :: rc
    end function tups_iteration

    function tups_calc_chi(ctx, sigma, sigma_h, mu, ngrain) bind(C, name="tups_calc_chi") result(rc)
      import :: c_ptr, c_int32_t, c_int, c_double
      type(c_ptr), value :: ctx
      type(c_ptr), value :: sigma, sigma_h
      real(c_double), value :: mu
      integer(c_int32_t), value :: ngrain
      integer(c_int) :: rc
    end function tups_calc_chi

    function tups_set_split(ctx, moduli, qbatch) bind(C, name="tups_set_split") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: moduli, qbatch
      integer(c_int) :: rc
    end function tups_set_split

    function tups_phir(ctx) bind(C, name="tups_phir") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function tups_phir

    function tups_build_m(ctx) bind(C, name="tups_build_m") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function tups_build_m

    function tups_pae(ctx) bind(C, name="tups_pae") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function tups_pae

    function tups_gamma_update(ctx) bind(C, name="tups_gamma_update") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function tups_gamma_update

    function tups_autotune(ctx, kernel, variant, ms) bind(C, name="tups_autotune") result(rc)
      import :: c_ptr, c_int32_t, c_float, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kernel
      integer(c_int32_t), intent(out) :: variant
      real(c_float), intent(out) :: ms
      integer(c_int) :: rc
    end function tups_autotune

    function tups_variant_name(kernel, variant) bind(C, name="tups_variant_name") result(p)
      import :: c_ptr, c_int32_t
      integer(c_int32_t), value :: kernel, variant
      type(c_ptr) :: p
    end function tups_variant_name

    function tups_set_variant(ctx, kernel, variant) bind(C, name="tups_set_variant") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kernel, variant
      integer(c_int) :: rc
    end function tups_set_variant

    function tups_variant_count(kernel) bind(C, name="tups_variant_count") result(n)
      import :: c_int32_t, c_int
      integer(c_int32_t), value :: kernel
      integer(c_int) :: n
    end function tups_variant_count

    function tups_sync(ctx) bind(C, name="tups_sync") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function tups_sync

    function tups_event_record(ctx, slot) bind(C, name="tups_event_record") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: slot
      integer(c_int) :: rc
    end function tups_event_record

    function tups_event_elapsed(ctx, a, b, ms) bind(C, name="tups_event_elapsed") result(rc)
      import :: c_ptr, c_int32_t, c_float, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: a, b
      real(c_float), intent(out) :: ms
      integer(c_int) :: rc
    end function tups_event_elapsed

    function tups_set_timing(ctx, on) bind(C, name="tups_set_timing") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
      integer(c_int) :: rc
    end function tups_set_timing

    function tups_get_timing(ctx, kernel, total_ms, launches) bind(C, name="tups_get_timing") result(rc)
      import :: c_ptr, c_int32_t, c_double, c_int64_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kernel
      real(c_double), intent(out) :: total_ms
      integer(c_int64_t), intent(out) :: launches
      integer(c_int) :: rc
    end function tups_get_timing

    function tups_check_finite(ctx, which, channel) bind(C, name="tups_check_finite") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which, channel
      integer(c_int) :: rc
    end function tups_check_finite

    function tups_synth_vertex(ctx, which, channel, seed, diag, sigma) &
        bind(C, name="tups_synth_vertex") result(rc)
      import :: c_ptr, c_int32_t, c_int64_t, c_double, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which, channel
      integer(c_int64_t), value :: seed
      real(c_double), value :: diag, sigma
      integer(c_int) :: rc
    end function tups_synth_vertex

    function tups_copy(ctx, which_dst, which_src, channel) bind(C, name="tups_copy") result(rc)
      import :: c_ptr, c_int32_t, c_int
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which_dst, which_src, channel
      integer(c_int) :: rc
    end function tups_copy

    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function c_strlen
  end interface

contains

  !> Stop with the library's message on a nonzero status.
  subroutine tups_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n, i
    character(len=512) :: text
    if (rc == TUPS_OK) return
    p = tups_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    text = ''
    do i = 1, min(n, len(text))
      text(i:i) = msg(i)
    end do
    write(*, *) 'ERROR - ', trim(where), ' failed (', rc, '): ', trim(text)
    stop 1
  end subroutine tups_check

end module tups_gpu_iface
