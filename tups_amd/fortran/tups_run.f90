!> A whole parquet run from the reference's parameter file, on the device.
!>
!> The steps of the reference's main program and execute_loop (SRC/loop.f90:29-249)
!> for its default settings: set_parameters reads the file (tups_read_prms); the
!> starting point of Initialization (SRC/parquet_util.f90:170-205,296-314,359-367):
!> Sigma = 0, Sigma_H = 0, mu = 0, G_r = F_r = 0 and the parquet approximation
!> Lambda_d = U, Lambda_m = -U, Lambda_s = 2U, Lambda_t = 0; the M tensors once;
!> then per iteration calc_Gkw_Grt's mu bisection, BSE, calc_PhiR, calculate_PAE,
!> calc_SDE (+ the use_U2 term), conv() and G = F - G until conv() reports
!> convergence or ite_max iterations have run.
!>
!> Usage: tups_run <prms file> [<Sigma file>]
!> Prints one line per iteration (mu, evaluations of the mu search, the Sigma
!> change and the vertex maxima of conv) and writes the final Sigma as records
!> "kx ky nu Re Im", the layout the reference reads as Sigma.init
!> (SRC/parquet_util.f90:183-194).  Settings that need input files or legacy
!> paths (DGA, use_sig_dmft, readinpick, oneshot, old, SDE_old, save_F) stop the
!> run with a message.
program tups_run
  use, intrinsic :: iso_c_binding
  use tups_gpu_iface
  implicit none

  integer, parameter :: dp = kind(1.0d0)
  character(len=512) :: fprm, fsig
  type(tups_prms) :: p
  type(tups_dims) :: d
  type(tups_sizes) :: sz
  type(tups_conv_info) :: info
  type(c_ptr) :: ctx
  complex(dp), allocatable :: lam(:, :, :), zero_v(:, :, :), sigma(:)
  real(dp) :: mu, lam_val(4)
  integer(c_int32_t) :: steps, r
  integer :: ite, u, kx, ky, nu, nt

  call get_command_argument(1, fprm)
  fsig = ''
  if (command_argument_count() >= 2) call get_command_argument(2, fsig)
  call tups_check(tups_read_prms(trim(fprm)//c_null_char, p), 'set_parameters')
  if (p%dga /= 0 .or. p%use_sig_dmft /= 0 .or. p%oneshot /= 0 .or. p%old /= 0 .or. p%sde_old /= 0 &
      .or. p%save_f /= 0) then
    write(*, *) 'tups_run: DGA, use_sig_dmft, oneshot, old, SDE_old and save_F must be .false.'
    stop 1
  end if
  call tups_check(tups_prms_dims(p, d), 'parameters')
  call tups_check(tups_create(d, ctx), 'tups_create')
  call tups_check(tups_query(ctx, sz), 'tups_query')
  write(*, '(a, f8.4, a, f8.4, a, f8.4, a, i4, a, i5, a, i3, a, i6)') ' beta =', p%beta, '  U =', p%u, &
      '  nParticle =', p%n_particle, '  Nx =', sz%nx, '  Nf =', sz%nf, '  Nl =', sz%nl, '  Nq =', sz%nq

  ! Lambda of the parquet approximation and G = F = 0
  lam_val = [p%u, -p%u, 2.0_dp * p%u, 0.0_dp]
  allocate(lam(sz%nf, sz%nf, sz%nw), zero_v(sz%nz, sz%nz, sz%nq))
  zero_v = (0.0_dp, 0.0_dp)
  do r = 0, 3
    lam = cmplx(lam_val(r + 1), 0.0_dp, dp)
    call tups_check(tups_upload(ctx, TUPS_ARR_LAMBDA, r, lam), 'Lambda')
    call tups_check(tups_upload(ctx, TUPS_ARR_GAMMA, r, zero_v), 'G')
    call tups_check(tups_upload(ctx, TUPS_ARR_F, r, zero_v), 'F')
  end do
  call tups_check(tups_build_m(ctx), 'calc_M')

  nt = sz%nx * sz%ny * sz%nf
  allocate(sigma(nt))
  sigma = (0.0_dp, 0.0_dp)
  call tups_check(tups_set_sigma(ctx, sigma, c_null_ptr, 0.0_dp, p%ngrain), 'Sigma')

  do ite = 0, p%ite_max - 1
    call tups_check(tups_iteration_full(ctx, TUPS_BSE_LADDER_FP64, 1_c_int32_t, p%f_range, p%u, p%use_u2, &
                                        p%n_particle), 'iteration')
    call tups_check(tups_get_mu(ctx, mu, steps), 'calc_Gkw_Grt')
    call tups_check(tups_conv(ctx, c_null_ptr, c_null_ptr, 1.0e-8_dp, 1.0e-8_dp, info), 'conv')
    write(*, '(a, i5, a, f14.10, a, i3, a, es12.4, a, es12.4, a, 3i4, a, 4es12.4)') ' iteration', ite, &
        '  mu', mu, '  evaluations', steps, '  Sigma avg err', info%sigma_avg, '  max', info%sigma_max, &
        '  at', info%kx, info%ky, info%nu, '  max|F|', info%f_max
    if (info%converged /= 0) then
      write(*, *) 'converged'
      exit
    end if
  end do

  call tups_check(tups_get_sigma(ctx, sigma), 'Sigma')
  if (len_trim(fsig) > 0) then
    open(newunit=u, file=trim(fsig), status='replace', action='write')
    do kx = 1, sz%nx
      do ky = 1, sz%ny
        do nu = 1, sz%nf
          write(u, '(3i6, 2es26.17)') kx, ky, nu, sigma(((kx - 1) * sz%ny + ky - 1) * sz%nf + nu)
        end do
      end do
    end do
    close(u)
  end if
  call tups_check(tups_destroy(ctx), 'tups_destroy')
end program tups_run
