!> Host state of the vertex hot path under the reference's names.
!>
!> The drop-in routines (tups_dropin.f90) are written against the module
!> variables the reference declares in parquet_ini (SRC/parquet_ini.f90:10-285),
!> parquet_EQMatrix (SRC/parquet_EQMatrix.f90:12-16) and the bubble routine
!> calc_chi (SRC/parquet_chi.f90:438).  Inside the reference those come from its
!> own modules; this module provides the same names so that the drop-in and its
!> driver build and run without the reference (which needs MPI and HDF5).
!> Integration replaces `use tups_host_state` by the reference's modules.
module tups_host_state
  use, intrinsic :: iso_c_binding
  implicit none

  integer, parameter :: dp = 8
  integer, parameter :: Ns = 8

  ! one rank holds every q (the q-split of SRC/parquet_ini.f90:336 collapses)
  integer :: id = 0, ntasks = 1, master = 0

  real(dp) :: beta, xU, f_damping
  integer :: Nx, Ny, Nf, Nl, Nz, Nb, NR, Nx_IBZ, NIBZ, Nred, Nc, wperTask

  type :: indxmap
    integer :: ix
    integer :: iy
    integer :: iw
  end type indxmap

  type(indxmap), allocatable :: Index_bosonic_IBZ(:)

  complex(dp), allocatable, dimension(:, :, :) :: F_d_LL, F_m_LL, F_s_LL, F_t_LL
  complex(dp), allocatable, dimension(:, :, :) :: G_d_LL, G_m_LL, G_s_LL, G_t_LL
  complex(dp), allocatable, dimension(:, :, :) :: L_d, L_m, L_s, L_t
  complex(dp), allocatable, dimension(:, :, :) :: PhiRd, PhiRm, PhiRs, PhiRt
  complex(dp), allocatable :: expSqR(:, :, :), expqSR(:, :, :)
  real(dp), allocatable :: FF(:, :, :), FF_inv(:, :, :, :)
  integer, allocatable :: signs(:), reps(:)
  complex(dp), allocatable, dimension(:, :, :, :, :, :) :: M_phbar_to_ph, M_pp_to_ph, &
                                                           M_ph_to_pp, M_phbar_to_pp

  ! bubbles handed to calc_chi by the driver (the reference computes them from G)
  complex(dp), allocatable :: chi_ph_all(:, :, :), chi_pp_all(:, :, :)

contains

  !> Sizes derived as in set_parameters, SRC/parquet_ini.f90:326-348, and the
  !> IBZ index map of SRC/parquet_util.f90:110-117.
  subroutine init_host_state(nx_in, nf_in, nl_in, nr_in, beta_in, f_in)
    integer, intent(in) :: nx_in, nf_in, nl_in, nr_in
    real(dp), intent(in) :: beta_in, f_in
    integer :: kx, ky, w, idx
    Nx = nx_in
    Ny = nx_in
    Nf = nf_in
    Nl = nl_in
    NR = nr_in
    beta = beta_in
    f_damping = f_in
    Nx_IBZ = (Nx + 2) / 2
    NIBZ = (Nx_IBZ * (Nx_IBZ + 1)) / 2
    Nred = NIBZ * Nf / 2
    Nc = Nx * Ny
    Nb = Nred / ntasks
    Nz = Nl * Nf
    wperTask = int((Nf - 1) / (2 * ntasks) + 1)
    allocate(Index_bosonic_IBZ(Nred))
    do kx = 1, Nx_IBZ
      do ky = 1, kx
        do w = 1, Nf / 2
          idx = (((kx - 1) * kx) / 2 + ky - 1) * Nf / 2 + w
          Index_bosonic_IBZ(idx) = indxmap(kx, ky, w)
        end do
      end do
    end do
    allocate(F_d_LL(Nz, Nz, Nb), F_m_LL(Nz, Nz, Nb), F_s_LL(Nz, Nz, Nb), F_t_LL(Nz, Nz, Nb))
    allocate(G_d_LL(Nz, Nz, Nb), G_m_LL(Nz, Nz, Nb), G_s_LL(Nz, Nz, Nb), G_t_LL(Nz, Nz, Nb))
    allocate(L_d(Nf, Nf, Nf / 2), L_m(Nf, Nf, Nf / 2), L_s(Nf, Nf, Nf / 2), L_t(Nf, Nf, Nf / 2))
    allocate(chi_ph_all(Nz, Nz, Nb), chi_pp_all(Nz, Nz, Nb))
    allocate(FF(Nx, Ny, Nl), FF_inv(Ns, Nx, Ny, Nl), signs(Nl), reps(Nl))
  end subroutine init_host_state

  !> list_index_IBZ, SRC/parquet_util.f90:512-520
  integer pure function list_index_IBZ(P) result(idx)
    type(indxmap), intent(in) :: P
    idx = (((P%ix - 1) * P%ix) / 2 + P%iy - 1) * Nf / 2 + P%iw
  end function list_index_IBZ

  !> Same signature as calc_chi (SRC/parquet_chi.f90:438); returns the driver's bubbles.
  subroutine calc_chi(map_q, chi_ph_LL, chi_pp_LL)
    type(indxmap), intent(in) :: map_q
    complex(dp), intent(out) :: chi_ph_LL(Nz, Nz)
    complex(dp), intent(out) :: chi_pp_LL(Nz, Nz)
    integer :: q
    q = list_index_IBZ(map_q)
    chi_ph_LL = chi_ph_all(:, :, q)
    chi_pp_LL = chi_pp_all(:, :, q)
  end subroutine calc_chi

end module tups_host_state
