!> One parquet iteration through the drop-in routines, in the order of
!> execute_loop (SRC/loop.f90:78-139, 219-222):
!>   BSE -> calculate_expSqR/expqSR -> calc_PhiR -> calc_M_* -> calculate_PAE
!>   -> G = F - G
!> then BSE again on the resident vertices (no upload), and once more after the
!> host halves F_r_LL and declares it with tups_host_modified.
!> Usage: tups_fortran_driver <input.bin> <output.bin>
!> input (stream): int32 nx, nf, nl, nr; real(8) beta, f_damping;
!>   complex(8) G_d,G_m,G_s,G_t, F_d..F_t (Nz,Nz,Nq), chi_ph, chi_pp (Nz,Nz,Nq),
!>   L_d..L_t (Nf,Nf,Nf/2)
!> output (stream): Phi_r (4), PhiR_r (4), M tensors (4, per q), F_r (4), G_r (4),
!>   Phi_r of the second BSE (4), Phi_r of the third BSE (4)
program tups_fortran_driver
  use, intrinsic :: iso_c_binding
  use tups_gpu_iface
  use tups_host_state
  use tups_dropin
  implicit none

  character(len=512) :: fin, fout
  integer(c_int32_t) :: hdr(4)
  real(dp) :: prm(2)
  type(tups_dims) :: d
  complex(dp), allocatable :: esq(:, :, :), eqs(:, :, :)
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  read(u) prm
  call init_host_state(int(hdr(1)), int(hdr(2)), int(hdr(3)), int(hdr(4)), prm(1), prm(2))
  read(u) G_d_LL, G_m_LL, G_s_LL, G_t_LL
  read(u) F_d_LL, F_m_LL, F_s_LL, F_t_LL
  read(u) chi_ph_all, chi_pp_all
  read(u) L_d, L_m, L_s, L_t
  close(u)

  ! the host's own form factors (FF_array in the reference's preprocessing)
  d = tups_dims(hdr(1), hdr(2), hdr(3), hdr(4), 0_c_int32_t, 0_c_int32_t, prm(1), prm(2))
  allocate(esq(NIBZ, NR, Ns), eqs(Nx * Ny, NR, Ns))
  call tups_check(tups_build_tables(d, FF, FF_inv, signs, reps, esq, eqs), 'tups_build_tables')

  call tups_gpu_init()
  call BSE(0)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) G_d_LL, G_m_LL, G_s_LL, G_t_LL
  call calculate_expSqR()
  call calculate_expqSR()
  call calc_PhiR(0)
  write(u) PhiRd, PhiRm, PhiRs, PhiRt
  call calc_M_phbar_to_ph()
  call calc_M_pp_to_ph()
  call calc_M_ph_to_pp()
  call calc_M_phbar_to_pp()
  write(u) M_phbar_to_ph, M_pp_to_ph, M_ph_to_pp, M_phbar_to_pp
  call calculate_PAE()
  write(u) F_d_LL, F_m_LL, F_s_LL, F_t_LL
  call update_vertices()
  write(u) G_d_LL, G_m_LL, G_s_LL, G_t_LL
  call BSE(1)
  write(u) G_d_LL, G_m_LL, G_s_LL, G_t_LL
  F_d_LL = 0.5d0 * F_d_LL
  F_m_LL = 0.5d0 * F_m_LL
  F_s_LL = 0.5d0 * F_s_LL
  F_t_LL = 0.5d0 * F_t_LL
  call tups_host_modified(.false., .true.)
  call BSE(2)
  write(u) G_d_LL, G_m_LL, G_s_LL, G_t_LL
  close(u)
  call tups_gpu_finalize()
  write(*, *) 'tups_fortran_driver: iteration done'
end program tups_fortran_driver
