! Fortran host harness: calls the reference-named forces / full_forces / realft / evolve_pi entry
! points (resolved to rpmd_mainz_amd/fortran/rpmd_hip.f90, i.e. the GPU library) exactly as the
! reference's callers do, after filling the reference COMMON blocks from an input file.
!
! usage: harness <input> <output>
!   input  (stream, native endian): int32 mode, na, nb, iopt; then the COMMON values
!          oo_eps oo_sig oo_gam rcut alpha ecut wm wh wrcut walpha rkmax (f64), kmax, iamcub (i32),
!          apot bpot alp alpb theta reoh sig (f64), nbdf1 nbdf2 mts (i32); box(3), z(na), r(3,na,nb)
!   modes 3 and 4 continue with: character(3) therm, int32 nsteps, ns (-1: no GLE), has_gp;
!          f64 ttau, dt, beta; int64 n_noise; mass(na), p(3,na,nb), dvdr(3,na,nb), dvdr2(3,na,nb),
!          noise(n_noise), A(ns+1,ns+1) as rows, gp(3,na,nb,ns+1) when has_gp
!   mode 0: forces(iopt) -> v, virial(9), dvdr(3,na,nb)
!   mode 1: full_forces  -> v, vew, voo, vint, vir(9), dvdr, dvdr2
!   mode 2: realft(r, 3*na, nb, iopt) -> r
!   mode 3: rpmd_set_thermostat, GLE set-up, rpmd_replay_noise, one evolve_pi step
!           -> v, vew, vlj, vint, vir(9), vir_lf(9), p, r, dvdr, dvdr2
!   mode 4: state_upload (seed iopt), nsteps device-resident steps, state_download
!           -> energies(4), vir(9), p, r, dvdr, dvdr2
!   mode 5: as mode 3 with nbaro = 1, after the mode-3 data: character(3) baro, f64 taub, patm, dv,
!           int32 n_u, f64 u(n_u), vir_lf(9); rpmd_set_barostat, rpmd_replay_uniforms when n_u > 0
!           -> as mode 3, then box(3), nacc, nmove (f64)
program harness
  use rpmd_hip
  implicit none
  integer mode, na, nb, iopt, kmax, iamcub_i, nbdf1, nbdf2, mts
  real(8) oo_eps, oo_sig, oo_gam, rcut, alpha, ecut, wm, wh, wrcut, walpha, rkmax
  real(8) apot, bpot, alp, alpb, theta, reoh, sig
  logical iamcub
  common /oo_param/ oo_eps, oo_sig, oo_gam, rcut
  common /ew_param/ alpha, ecut, wm, wh
  common /ew_setup/ wrcut, walpha, rkmax, kmax
  common /intra_param/ apot, bpot, alp, alpb
  common /geometry/ theta, reoh
  common /symmetry/ iamcub
  common /correct/ sig
  common /beaddiabatic/ nbdf1, nbdf2
  common /multiple_ts/ mts
  real(8), allocatable :: r(:,:,:), dvdr(:,:,:), dvdr2(:,:,:), z(:), p(:,:,:), mass(:), noise(:), a(:,:)
  real(8), allocatable :: gp(:,:,:,:)
  real(8) box(3), v, vew, voo, vint, vir(3,3), vir_lf(3,3), ttau, dt, beta, en(4)
  character(len=3) therm, baro
  integer nsteps, ns, has_gp, i, n_u, nacc, nmove
  real(8) taub, patm, dv
  real(8), allocatable :: u(:)
  integer(8) n_noise
  character(len=512) fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) mode, na, nb, iopt
  read(10) oo_eps, oo_sig, oo_gam, rcut, alpha, ecut, wm, wh, wrcut, walpha, rkmax
  read(10) kmax, iamcub_i
  read(10) apot, bpot, alp, alpb, theta, reoh, sig
  read(10) nbdf1, nbdf2, mts
  iamcub = iamcub_i .ne. 0
  allocate(r(3,na,nb), dvdr(3,na,nb), dvdr2(3,na,nb), z(na), p(3,na,nb), mass(na))
  read(10) box, z, r
  if (mode .ge. 3) then
     read(10) therm
     read(10) nsteps, ns, has_gp
     read(10) ttau, dt, beta
     read(10) n_noise
     allocate(noise(max(n_noise, 1_8)))
     read(10) mass, p, dvdr, dvdr2
     if (n_noise .gt. 0) read(10) noise(1:n_noise)
     if (ns .ge. 0) then
        allocate(a(ns+1, ns+1))
        do i = 1, ns + 1
           read(10) a(i, :)
        enddo
        if (has_gp .ne. 0) then
           allocate(gp(3, na, nb, ns + 1))
           read(10) gp
        endif
     endif
     if (mode .eq. 5) then
        read(10) baro
        read(10) taub, patm, dv
        read(10) n_u
        allocate(u(max(n_u, 1)))
        if (n_u .gt. 0) read(10) u(1:n_u)
        read(10) vir_lf
     endif
  endif
  close(10)
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  if (mode .eq. 0) then
     call forces(r, v, dvdr, nb, na, box, z, vir, iopt)
     write(11) v, vir, dvdr
  else if (mode .eq. 1) then
     call full_forces(r, na, nb, v, vew, voo, vint, vir, z, box, dvdr, dvdr2)
     write(11) v, vew, voo, vint, vir, dvdr, dvdr2
  else if (mode .eq. 2) then
     call realft(r, 3*na, nb, iopt)
     write(11) r
  else if (mode .eq. 3) then
     call rpmd_set_thermostat(therm, ttau)
     if (ns .ge. 0) then
        if (has_gp .ne. 0) then
           call rpmd_gle_setup(ns, a, na, nb, 0.5d0 * dt, beta, gp)
        else
           call rpmd_gle_setup(ns, a, na, nb, 0.5d0 * dt, beta, seed=iopt)
        endif
     endif
     if (n_noise .gt. 0) call rpmd_replay_noise(noise(1:n_noise))
     call evolve_pi(p, r, v, vew, voo, vint, dvdr, dvdr2, dt, mass, na, nb, box, z, beta, vir, vir_lf, iopt, 0)
     write(11) v, vew, voo, vint, vir, vir_lf, p, r, dvdr, dvdr2
  else if (mode .eq. 5) then
     call rpmd_set_thermostat(therm, ttau)
     call rpmd_set_barostat(baro, taub, patm, dv)
     if (n_noise .gt. 0) call rpmd_replay_noise(noise(1:n_noise))
     if (n_u .gt. 0) call rpmd_replay_uniforms(u(1:n_u))
     call evolve_pi(p, r, v, vew, voo, vint, dvdr, dvdr2, dt, mass, na, nb, box, z, beta, vir, vir_lf, iopt, 1)
     call rpmd_barostat_counts(nacc, nmove)
     write(11) v, vew, voo, vint, vir, vir_lf, p, r, dvdr, dvdr2, box, dble(nacc), dble(nmove)
  else
     ! the device-resident API through its bind(C) interfaces
     call rpmd_set_thermostat(therm, ttau)
     call rpmd_bind(na, nb)
     if (ns .ge. 0) call rpmd_gle_setup(ns, a, na, nb, 0.5d0 * dt, beta, seed=iopt)
     call rpmd_check(c_state_upload(rpmd_ctx, p, r, mass, z, int(na, c_int), int(nb, c_int), box, beta, dt, &
                                    int(iopt, c_long_long)))
     call rpmd_check(c_use_graph(rpmd_ctx, 1_c_int))
     call rpmd_check(c_run_steps(rpmd_ctx, int(nsteps, c_int)))
     call rpmd_check(c_state_download(rpmd_ctx, p, r, dvdr, dvdr2, en, vir))
     write(11) en, vir, p, r, dvdr, dvdr2
  endif
  close(11)
end program harness
