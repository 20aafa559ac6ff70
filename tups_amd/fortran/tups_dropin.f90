!> Drop-in replacements of the reference's hot-path routines, backed by the
!> MI355X library through ISO_C_BINDING.
!>
!> Same names and argument lists as the reference:
!>   BSE(ite)                 SRC/parquet_BSE.f90:19
!>   calculate_expSqR()       SRC/parquet_PhiR.f90:360
!>   calculate_expqSR()       SRC/parquet_PhiR.f90:400
!>   calc_PhiR(ite)           SRC/parquet_PhiR.f90:35
!>   calc_M_phbar_to_ph() ... SRC/parquet_EQMatrix.f90:21,126,198,274
!>   calculate_PAE()          SRC/parquet_EQOuter.f90:16
!> plus tups_gpu_init / tups_gpu_finalize and update_vertices (the
!> G = F - G step of SRC/loop.f90:219-222 on the device).
!>
!> Residency: the vertex arrays stay in HBM between calls; the host arrays are
!> kept as mirrors (G_r_LL after BSE holds Phi for the host-side SDE,
!> SRC/parquet_SDE.f90:523-566; F_r_LL after calculate_PAE; G_r_LL after
!> update_vertices).  BSE uploads G_r_LL and F_r_LL only when the device copy
!> may differ from the host array: before the first BSE, and after the host
!> declared a change with tups_host_modified.  In execute_loop
!> (SRC/loop.f90:78-222) the host writes these arrays only in the G = F - G
!> step, which update_vertices does on the device; a host that keeps that step
!> (or changes the arrays elsewhere) calls tups_host_modified(.true., .false.)
!> after it, or sets tups_trust_residency = .false. to upload on every BSE.
module tups_dropin
  use, intrinsic :: iso_c_binding
  use tups_gpu_iface
  use tups_host_state
  implicit none

  type(c_ptr), save :: tups_h = c_null_ptr
  logical, save :: m_built = .false.
  logical, save :: lambda_on_device = .false.
  ! arithmetic of BSE(ite): TUPS_BSE_LADDER_FP64 (default) or TUPS_BSE_LADDER_SPLIT
  integer(c_int32_t), save :: tups_bse_mode = TUPS_BSE_LADDER_FP64
  ! .false.: BSE uploads G_r_LL and F_r_LL on every call
  logical, save :: tups_trust_residency = .true.
  ! the device copy of G_r_LL (its Gamma slot) / F_r_LL equals the host array
  logical, save :: g_on_device = .false.
  logical, save :: f_on_device = .false.

contains

  subroutine tups_gpu_init()
    type(tups_dims) :: d
    real(dp), allocatable :: ff_dev(:, :, :), ffi_dev(:, :, :, :)
    integer(c_int32_t), allocatable :: sg(:), rp(:)
    complex(dp), allocatable :: esq(:, :, :), eqs(:, :, :)
    if (ntasks /= 1) then
      write(*, *) 'ERROR - the GPU hot path holds every q of the box on one device (ntasks = 1)'
      stop 1
    end if
    d = tups_dims(int(Nx, c_int32_t), int(Nf, c_int32_t), int(Nl, c_int32_t), int(NR, c_int32_t), &
                  0_c_int32_t, 0_c_int32_t, real(beta, c_double), real(f_damping, c_double))
    call tups_check(tups_create(d, tups_h), 'tups_create')
    ! the library's basis must be the one the host uses (FF_array, SRC/parquet_formfactors.f90:11-90)
    allocate(ff_dev(Nx, Ny, Nl), ffi_dev(Ns, Nx, Ny, Nl), sg(Nl), rp(Nl))
    allocate(esq(NIBZ, NR, Ns), eqs(Nx * Ny, NR, Ns))
    call tups_check(tups_get_tables(tups_h, ff_dev, ffi_dev, sg, rp, esq, eqs), 'tups_get_tables')
    if (maxval(abs(FF - ff_dev)) > 1.0d-13 .or. maxval(abs(FF_inv - ffi_dev)) > 1.0d-13 .or. &
        any(signs /= sg) .or. any(reps /= rp)) then
      write(*, *) 'ERROR - host form factors differ from the library basis'
      stop 1
    end if
    m_built = .false.
    lambda_on_device = .false.
    g_on_device = .false.
    f_on_device = .false.
  end subroutine tups_gpu_init

  !> The host changed G_r_LL (gamma) and/or F_r_LL (f) since the library last
  !> saw them: the next BSE uploads them.
  subroutine tups_host_modified(gamma, f)
    logical, intent(in) :: gamma, f
    if (gamma) g_on_device = .false.
    if (f) f_on_device = .false.
  end subroutine tups_host_modified

  subroutine tups_gpu_finalize()
    if (c_associated(tups_h)) call tups_check(tups_destroy(tups_h), 'tups_destroy')
    tups_h = c_null_ptr
  end subroutine tups_gpu_finalize

  !> BSE(ite): Phi_r = f c_r (alpha Gamma chi F + F alpha chi Gamma) + (1-f)(F - Gamma)
  !> into G_r_LL, SRC/parquet_BSE.f90:62-132.
  subroutine BSE(ite)
    integer, intent(in) :: ite
    integer :: idx_q
    type(indxmap) :: map_q
    complex(dp), allocatable :: chi_ph_LL(:, :), chi_pp_LL(:, :)
    allocate(chi_ph_LL(Nz, Nz), chi_pp_LL(Nz, Nz))
    do idx_q = 1, Nb
      map_q = Index_Bosonic_IBZ(idx_q + id * Nb)
      call calc_chi(map_q, chi_ph_LL, chi_pp_LL)
      call tups_check(tups_upload_chi_dense(tups_h, 0_c_int32_t, int(idx_q - 1, c_int32_t), 1_c_int32_t, &
                                            chi_ph_LL), 'upload chi_ph')
      call tups_check(tups_upload_chi_dense(tups_h, 1_c_int32_t, int(idx_q - 1, c_int32_t), 1_c_int32_t, &
                                            chi_pp_LL), 'upload chi_pp')
    end do
    if (.not. (g_on_device .and. tups_trust_residency)) then
      call put(TUPS_ARR_GAMMA, TUPS_CH_D, G_d_LL)
      call put(TUPS_ARR_GAMMA, TUPS_CH_M, G_m_LL)
      call put(TUPS_ARR_GAMMA, TUPS_CH_S, G_s_LL)
      call put(TUPS_ARR_GAMMA, TUPS_CH_T, G_t_LL)
    end if
    if (.not. (f_on_device .and. tups_trust_residency)) then
      call put(TUPS_ARR_F, TUPS_CH_D, F_d_LL)
      call put(TUPS_ARR_F, TUPS_CH_M, F_m_LL)
      call put(TUPS_ARR_F, TUPS_CH_S, F_s_LL)
      call put(TUPS_ARR_F, TUPS_CH_T, F_t_LL)
      f_on_device = .true.
    end if
    call tups_check(tups_bse(tups_h, tups_bse_mode), 'BSE')
    call get(TUPS_ARR_GAMMA, TUPS_CH_D, G_d_LL)
    call get(TUPS_ARR_GAMMA, TUPS_CH_M, G_m_LL)
    call get(TUPS_ARR_GAMMA, TUPS_CH_S, G_s_LL)
    call get(TUPS_ARR_GAMMA, TUPS_CH_T, G_t_LL)
    g_on_device = .true.  ! the Gamma slot holds Phi, as G_r_LL now does
    if (ite < 0) stop 1
  end subroutine BSE

  !> calculate_expSqR: e^{i (SQ).R} on the star of Q (library table), SRC/parquet_PhiR.f90:360-395
  subroutine calculate_expSqR()
    real(dp), allocatable :: f1(:, :, :), f2(:, :, :, :)
    integer(c_int32_t), allocatable :: s1(:), s2(:)
    complex(dp), allocatable :: tmp(:, :, :)
    if (.not. allocated(expSqR)) allocate(expSqR(NIBZ, NR, Ns))
    allocate(f1(Nx, Ny, Nl), f2(Ns, Nx, Ny, Nl), s1(Nl), s2(Nl), tmp(Nx * Ny, NR, Ns))
    call tups_check(tups_get_tables(tups_h, f1, f2, s1, s2, expSqR, tmp), 'calculate_expSqR')
  end subroutine calculate_expSqR

  !> calculate_expqSR: e^{-i k.(S R)} on the star of R (library table), SRC/parquet_PhiR.f90:400-435
  subroutine calculate_expqSR()
    real(dp), allocatable :: f1(:, :, :), f2(:, :, :, :)
    integer(c_int32_t), allocatable :: s1(:), s2(:)
    complex(dp), allocatable :: tmp(:, :, :)
    if (.not. allocated(expqSR)) allocate(expqSR(Nx * Ny, NR, Ns))
    allocate(f1(Nx, Ny, Nl), f2(Ns, Nx, Ny, Nl), s1(Nl), s2(Nl), tmp(NIBZ, NR, Ns))
    call tups_check(tups_get_tables(tups_h, f1, f2, s1, s2, tmp, expqSR), 'calculate_expqSR')
  end subroutine calculate_expqSR

  !> calc_PhiR(ite): PhiR_r(Nz*Nz, w, R) from the resident Phi_r, SRC/parquet_PhiR.f90:35-109
  subroutine calc_PhiR(ite)
    integer, intent(in) :: ite
    if (.not. allocated(PhiRd)) allocate(PhiRd(Nz * Nz, wperTask, NR))
    if (.not. allocated(PhiRm)) allocate(PhiRm(Nz * Nz, wperTask, NR))
    if (.not. allocated(PhiRs)) allocate(PhiRs(Nz * Nz, wperTask, NR))
    if (.not. allocated(PhiRt)) allocate(PhiRt(Nz * Nz, wperTask, NR))
    call tups_check(tups_phir(tups_h), 'calc_PhiR')
    call get(TUPS_ARR_PHIR, TUPS_CH_D, PhiRd)
    call get(TUPS_ARR_PHIR, TUPS_CH_M, PhiRm)
    call get(TUPS_ARR_PHIR, TUPS_CH_S, PhiRs)
    call get(TUPS_ARR_PHIR, TUPS_CH_T, PhiRt)
    if (ite < 0) stop 1
  end subroutine calc_PhiR

  subroutine ensure_m()
    if (.not. m_built) then
      call tups_check(tups_build_m(tups_h), 'calc_M')
      m_built = .true.
    end if
  end subroutine ensure_m

  !> Expand the per-momentum device tensor to the reference's per-q layout M(.,.,.,.,Nb,NR).
  subroutine fetch_m(which, m)
    integer(c_int32_t), intent(in) :: which
    complex(dp), allocatable, intent(inout) :: m(:, :, :, :, :, :)
    complex(dp), allocatable :: dev(:, :, :, :, :, :)
    integer :: q, mom, r
    allocate(dev(Nl, Nl, Nl, Nl, NIBZ, NR))
    call tups_check(tups_download(tups_h, TUPS_ARR_M, which, dev), 'download M')
    if (.not. allocated(m)) allocate(m(Nl, Nl, Nl, Nl, Nb, NR))
    do r = 1, NR
      do q = 1, Nb
        mom = (q + id * Nb - 1) / (Nf / 2) + 1
        m(:, :, :, :, q, r) = dev(:, :, :, :, mom, r)
      end do
    end do
  end subroutine fetch_m

  subroutine calc_M_phbar_to_ph()
    call ensure_m()
    call fetch_m(0_c_int32_t, M_phbar_to_ph)
  end subroutine calc_M_phbar_to_ph

  subroutine calc_M_pp_to_ph()
    call ensure_m()
    call fetch_m(1_c_int32_t, M_pp_to_ph)
  end subroutine calc_M_pp_to_ph

  subroutine calc_M_ph_to_pp()
    call ensure_m()
    call fetch_m(2_c_int32_t, M_ph_to_pp)
  end subroutine calc_M_ph_to_pp

  subroutine calc_M_phbar_to_pp()
    call ensure_m()
    call fetch_m(3_c_int32_t, M_phbar_to_pp)
  end subroutine calc_M_phbar_to_pp

  !> calculate_PAE: F_r = Phi_r + L_r*Nc + crossing, SRC/parquet_EQOuter.f90:16-118
  subroutine calculate_PAE()
    if (.not. lambda_on_device) then
      call put(TUPS_ARR_LAMBDA, TUPS_CH_D, L_d)
      call put(TUPS_ARR_LAMBDA, TUPS_CH_M, L_m)
      call put(TUPS_ARR_LAMBDA, TUPS_CH_S, L_s)
      call put(TUPS_ARR_LAMBDA, TUPS_CH_T, L_t)
      lambda_on_device = .true.
    end if
    call ensure_m()
    call tups_check(tups_pae(tups_h), 'calculate_PAE')
    call get(TUPS_ARR_F, TUPS_CH_D, F_d_LL)
    call get(TUPS_ARR_F, TUPS_CH_M, F_m_LL)
    call get(TUPS_ARR_F, TUPS_CH_S, F_s_LL)
    call get(TUPS_ARR_F, TUPS_CH_T, F_t_LL)
    f_on_device = .true.
  end subroutine calculate_PAE

  !> G_r = F_r - G_r on the device, SRC/loop.f90:219-222
  subroutine update_vertices()
    call tups_check(tups_gamma_update(tups_h), 'update_vertices')
    call get(TUPS_ARR_GAMMA, TUPS_CH_D, G_d_LL)
    call get(TUPS_ARR_GAMMA, TUPS_CH_M, G_m_LL)
    call get(TUPS_ARR_GAMMA, TUPS_CH_S, G_s_LL)
    call get(TUPS_ARR_GAMMA, TUPS_CH_T, G_t_LL)
    g_on_device = .true.
  end subroutine update_vertices

  subroutine put(which, ch, a)
    integer(c_int32_t), intent(in) :: which, ch
    complex(dp), intent(in) :: a(*)
    call tups_check(tups_upload(tups_h, which, ch, a), 'upload')
  end subroutine put

  subroutine get(which, ch, a)
    integer(c_int32_t), intent(in) :: which, ch
    complex(dp), intent(out) :: a(*)
    call tups_check(tups_download(tups_h, which, ch, a), 'download')
  end subroutine get

end module tups_dropin
