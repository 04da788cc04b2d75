! rpmd_hip.f90 -- ISO_C_BINDING layer between the reference's Fortran host and librpmd_mi355x.so.
!
! Link this file in place of the reference's md_forces.f90 (forces, full_forces), fourier.f90
! (realft), evolve_pi.f90 (evolve_pi) and evolve_rigid.f90's evolve_rig_pi: the external subroutines
! below keep the reference's names and dummy-argument lists
!   forces(r,v,dvdr,nb,na,boxlxyz,z,virial,iopt)                src/md_forces.f90:1
!   full_forces(r,na,nb,v,vew,voo,vint,vir,z,boxlxyz,dvdr,dvdr2) src/md_forces.f90:305
!   realft(data,m,n,mode)                                       src/fourier.f90:16
!   evolve_pi(p,r,v,vew,vlj,vint,dvdr,dvdr2,dt,mass,na,nb,boxlxyz,z,beta,vir,vir_lf,irun,nbaro)
!                                                               src/evolve_pi.f90:1
!   evolve_rig_pi(p,r,v,v_lf,v_hf,dvdr,dvdr2,dt,mass,na,nb,boxlxyz,z,beta,vir,vir_lf,irun,nbaro)
!                                                               src/evolve_rigid.f90:100
!   run_driver(p,r,dvdr,dvdr2,na,nb,boxlxyz,z,beta,dt,mass,irun,itst,pt,pb,print)
!                                                               src/ipi_driver.f90:1 (with sockets.c)
! so every existing call site compiles and links unchanged.  Model parameters are read from the
! same COMMON blocks the reference routines read (src/main.f90:53-72) on every call; the
! thermostat (module thermostat's therm and ttau, src/module_thermostat.f90:3-4) is handed over
! once with rpmd_set_thermostat, since that module belongs to the host.  A context for the largest
! (na, nb) seen is created on first use.  A non-zero status from the library stops the program with
! the reference's message, as the reference's STOP statements do.
module rpmd_hip
  use iso_c_binding
  implicit none

  type, bind(C) :: rpmd_params_t
     real(c_double) :: oo_eps, oo_sig, oo_gam, rcut
     real(c_double) :: alpha, ecut, wm, wh
     real(c_double) :: wrcut, walpha, rkmax
     integer(c_int) :: kmax, iamcub
     real(c_double) :: apot, bpot, alp, alpb
     real(c_double) :: theta, reoh
     real(c_double) :: sig
     integer(c_int) :: nbdf1, nbdf2, mts
     integer(c_int) :: intra_model, therm
     real(c_double) :: ttau
     integer(c_int) :: rigid
     real(c_double) :: rat_tol
  end type rpmd_params_t

  interface
     integer(c_int) function c_ctx_create(prm, na, nb_max, precision, ctx) bind(C, name='rpmd_ctx_create')
       import :: c_int, c_ptr, rpmd_params_t
       type(rpmd_params_t), intent(in) :: prm
       integer(c_int), value :: na, nb_max, precision
       type(c_ptr), intent(out) :: ctx
     end function c_ctx_create
     integer(c_int) function c_ctx_destroy(ctx) bind(C, name='rpmd_ctx_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function c_ctx_destroy
     integer(c_int) function c_ctx_set_params(ctx, prm) bind(C, name='rpmd_ctx_set_params')
       import :: c_int, c_ptr, rpmd_params_t
       type(c_ptr), value :: ctx
       type(rpmd_params_t), intent(in) :: prm
     end function c_ctx_set_params
     integer(c_int) function c_forces(ctx, r, nb, na, box, z, iopt, v, dvdr, virial) bind(C, name='rpmd_forces')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: r(*), box(3), z(*)
       integer(c_int), value :: nb, na, iopt
       real(c_double), intent(out) :: v, dvdr(*), virial(9)
     end function c_forces
     integer(c_int) function c_full_forces(ctx, r, na, nb, v, vew, voo, vint, vir, z, box, dvdr, dvdr2) &
          bind(C, name='rpmd_full_forces')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: r(*), z(*), box(3)
       integer(c_int), value :: na, nb
       real(c_double), intent(out) :: v, vew, voo, vint, vir(9), dvdr(*), dvdr2(*)
     end function c_full_forces
     integer(c_int) function c_realft(ctx, data, m, n, mode) bind(C, name='rpmd_realft')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: data(*)
       integer(c_int), value :: m, n, mode
     end function c_realft
     ! the whole step, src/evolve_pi.f90:1 (noise: c_null_ptr draws on the device)
     integer(c_int) function c_evolve_pi(ctx, p, r, v, vew, vlj, vint, dvdr, dvdr2, dt, mass, na, nb, box, z, &
          beta, vir, vir_lf, noise, n_noise) bind(C, name='rpmd_evolve_pi')
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       real(c_double) :: p(*), r(*), v, vew, vlj, vint, dvdr(*), dvdr2(*), vir(9), vir_lf(9)
       real(c_double), value :: dt, beta
       real(c_double), intent(in) :: mass(*), box(3), z(*)
       integer(c_int), value :: na, nb
       type(c_ptr), value :: noise
       integer(c_long_long), value :: n_noise
     end function c_evolve_pi
     ! evolve_pi with nbaro = 1 (box and vir_lf in/out), src/evolve_pi.f90:163-189
     integer(c_int) function c_evolve_pi_npt(ctx, p, r, v, vew, vlj, vint, dvdr, dvdr2, dt, mass, na, nb, box, z, &
          beta, vir, vir_lf, noise, n_noise, uniforms, n_uniforms) bind(C, name='rpmd_evolve_pi_npt')
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       real(c_double) :: p(*), r(*), v, vew, vlj, vint, dvdr(*), dvdr2(*), vir(9), vir_lf(9), box(3)
       real(c_double), value :: dt, beta
       real(c_double), intent(in) :: mass(*), z(*)
       integer(c_int), value :: na, nb
       type(c_ptr), value :: noise, uniforms
       integer(c_long_long), value :: n_noise
       integer(c_int), value :: n_uniforms
     end function c_evolve_pi_npt
     ! evolve_rig_pi with nbaro = 1 (box in/out), src/evolve_rigid.f90:203-220
     integer(c_int) function c_evolve_rig_pi_npt(ctx, p, r, v, v_lf, v_hf, dvdr, dvdr2, dt, mass, na, nb, box, z, &
          beta, vir, vir_lf, noise, n_noise, uniforms, n_uniforms) bind(C, name='rpmd_evolve_rig_pi_npt')
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       real(c_double) :: p(*), r(*), v, v_lf, v_hf, dvdr(*), dvdr2(*), vir(9), vir_lf(9), box(3)
       real(c_double), value :: dt, beta
       real(c_double), intent(in) :: mass(*), z(*)
       integer(c_int), value :: na, nb
       type(c_ptr), value :: noise, uniforms
       integer(c_long_long), value :: n_noise
       integer(c_int), value :: n_uniforms
     end function c_evolve_rig_pi_npt
     ! module barostat's settings and MC counters
     integer(c_int) function c_set_barostat(ctx, baro, taub, patm, dv) bind(C, name='rpmd_set_barostat')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: baro
       real(c_double), value :: taub, patm, dv
     end function c_set_barostat
     integer(c_int) function c_barostat_counts(ctx, nacc, nmove) bind(C, name='rpmd_barostat_counts')
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long), intent(out) :: nacc, nmove
     end function c_barostat_counts
     ! the rigid step, src/evolve_rigid.f90:100 (vir in/out as there)
     integer(c_int) function c_evolve_rig_pi(ctx, p, r, v, v_lf, v_hf, dvdr, dvdr2, dt, mass, na, nb, box, z, &
          beta, vir, vir_lf, noise, n_noise) bind(C, name='rpmd_evolve_rig_pi')
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       real(c_double) :: p(*), r(*), v, v_lf, v_hf, dvdr(*), dvdr2(*), vir(9), vir_lf(9)
       real(c_double), value :: dt, beta
       real(c_double), intent(in) :: mass(*), box(3), z(*)
       integer(c_int), value :: na, nb
       type(c_ptr), value :: noise
       integer(c_long_long), value :: n_noise
     end function c_evolve_rig_pi
     ! therm_gle_init, src/module_gle.f90:146 (A in the GLE-A row order; gp: c_null_ptr draws)
     integer(c_int) function c_gle_init(ctx, ns, a, na, nb, dt, beta, gp, seed) bind(C, name='rpmd_gle_init')
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       integer(c_int), value :: ns, na, nb
       real(c_double), intent(in) :: a(*)
       real(c_double), value :: dt, beta
       type(c_ptr), value :: gp
       integer(c_long_long), value :: seed
     end function c_gle_init
     type(c_ptr) function c_strerror(code) bind(C, name='rpmd_strerror')
       import :: c_int, c_ptr
       integer(c_int), value :: code
     end function c_strerror
     ! host-side Ewald set-up, src/setup_options.f90:104-189
     integer(c_int) function c_setup_ewald(na, box, prm) bind(C, name='rpmd_setup_ewald')
       import :: c_int, c_double, rpmd_params_t
       integer(c_int), value :: na
       real(c_double), intent(in) :: box(3)
       type(rpmd_params_t), intent(inout) :: prm
     end function c_setup_ewald
     integer(c_int) function c_ctx_set_precision(ctx, precision) bind(C, name='rpmd_ctx_set_precision')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: precision
     end function c_ctx_set_precision
     ! sample_rpmd, src/module_thermostat.f90:199 (noise: c_null_ptr draws on the device from seed)
     integer(c_int) function c_sample_rpmd(ctx, p, na, nb, mass, beta, noise, n_noise, seed) &
          bind(C, name='rpmd_sample_rpmd')
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: p(*)
       integer(c_int), value :: na, nb
       real(c_double), intent(in) :: mass(*)
       real(c_double), value :: beta
       type(c_ptr), value :: noise
       integer(c_long_long), value :: n_noise, seed
     end function c_sample_rpmd
     ! the free ring-polymer update of evolve_pi (normal-mode p, r), src/evolve_pi.f90:109-125
     integer(c_int) function c_free_ring(ctx, p, r, mass, na, nb, dt, beta) bind(C, name='rpmd_free_ring')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: p(*), r(*)
       real(c_double), intent(in) :: mass(*)
       integer(c_int), value :: na, nb
       real(c_double), value :: dt, beta
     end function c_free_ring
     ! realft without a context or device allocation
     integer(c_int) function c_realft_noalloc(data, m, n, mode) bind(C, name='rpmd_realft_noalloc')
       import :: c_int, c_double
       real(c_double), intent(inout) :: data(*)
       integer(c_int), value :: m, n, mode
     end function c_realft_noalloc
     ! device-resident stepping: upload once, run steps in HBM, read back when needed
     integer(c_int) function c_state_upload(ctx, p, r, mass, z, na, nb, box, beta, dt, seed) &
          bind(C, name='rpmd_state_upload')
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: p(*), r(*), mass(*), z(*), box(3)
       integer(c_int), value :: na, nb
       real(c_double), value :: beta, dt
       integer(c_long_long), value :: seed
     end function c_state_upload
     integer(c_int) function c_run_steps(ctx, nsteps) bind(C, name='rpmd_run_steps')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function c_run_steps
     integer(c_int) function c_state_download(ctx, p, r, dvdr, dvdr2, energies, vir) bind(C, name='rpmd_state_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: p(*), r(*), dvdr(*), dvdr2(*), energies(4), vir(9)
     end function c_state_download
     integer(c_int) function c_use_graph(ctx, enable) bind(C, name='rpmd_use_graph')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: enable
     end function c_use_graph
     integer(c_int) function c_ctx_set_streams(ctx, streams) bind(C, name='rpmd_ctx_set_streams')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: streams
     end function c_ctx_set_streams
     ! replica batches: contexts with uploaded states advanced together
     integer(c_int) function c_batch_create(ctxs, n, batch) bind(C, name='rpmd_batch_create')
       import :: c_int, c_ptr
       type(c_ptr), intent(in) :: ctxs(*)
       integer(c_int), value :: n
       type(c_ptr), intent(out) :: batch
     end function c_batch_create
     integer(c_int) function c_batch_run_steps(batch, nsteps) bind(C, name='rpmd_batch_run_steps')
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), value :: nsteps
     end function c_batch_run_steps
     integer(c_int) function c_batch_destroy(batch) bind(C, name='rpmd_batch_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
     end function c_batch_destroy
     ! the autotuners' choices of a batch
     integer(c_int) function c_batch_tuning(batch, batched, ms_batched, ms_separate, pair_cap, pair_beside) &
         bind(C, name='rpmd_batch_tuning')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: batch
       integer(c_int), intent(out) :: batched, pair_cap, pair_beside
       real(c_double), intent(out) :: ms_batched, ms_separate
     end function c_batch_tuning
     ! estimators of the device-resident state (RPMD_EST_* order, 11 values)
     integer(c_int) function c_estimators(ctx, est) bind(C, name='rpmd_estimators')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: est(11)
     end function c_estimators
     ! the i-PI client (run_driver)
     integer(c_int) function c_ipi_run(ctx, address, z, na, served) bind(C, name='rpmd_ipi_run')
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), dimension(*), intent(in) :: address
       real(c_double), intent(in) :: z(*)
       integer(c_int), value :: na
       integer(c_int), intent(out) :: served
     end function c_ipi_run
  end interface

  type(c_ptr), save :: rpmd_ctx = c_null_ptr
  integer, save :: ctx_na = 0, ctx_nb = 0
  ! RPMD_PREC_FP64 = 0, RPMD_PREC_MIXED = 1; set before the first call to choose the mode
  integer, save :: rpmd_precision = 0
  ! the thermostat of the step entry points (RPMD_THERM_* and ttau in a.u. time), as the host's
  ! module thermostat holds it; set with rpmd_set_thermostat
  integer, save :: rpmd_therm = 0
  real(8), save :: rpmd_ttau = 0.d0
  ! RATTLE tolerance of the rigid step (<= 0: the reference's 1e-6)
  real(8), save :: rpmd_rattle_tol = 0.d0
  ! replayed Gaussian deviates for the step entry points (rpmd_replay_noise); none: the device's
  ! own counter-based stream
  real(c_double), allocatable, target, save :: rpmd_noise(:)
  ! the barostat of evolve_pi steps with nbaro = 1 (module barostat: RPMD_BARO_*, taub, patm, dv),
  ! set with rpmd_set_barostat; replayed MC draws (rpmd_replay_uniforms), else the library's stream
  integer, save :: rpmd_baro = 0
  real(8), save :: rpmd_taub = 0.d0, rpmd_patm = 0.d0, rpmd_dv = 0.d0
  real(c_double), allocatable, target, save :: rpmd_uniforms(:)

contains

  subroutine rpmd_fill_params(prm)
    ! the reference COMMON blocks, as declared in src/main.f90:53-72
    type(rpmd_params_t), intent(out) :: prm
    real(8) oo_eps, oo_sig, oo_gam, rcut, alpha, ecut, wm, wh, wrcut, walpha, rkmax
    real(8) apot, bpot, alp, alpb, theta, reoh, sig
    integer kmax, nbdf1, nbdf2, mts
    logical iamcub, iamrigid
    common /oo_param/ oo_eps, oo_sig, oo_gam, rcut
    common /ew_param/ alpha, ecut, wm, wh
    common /ew_setup/ wrcut, walpha, rkmax, kmax
    common /intra_param/ apot, bpot, alp, alpb
    common /geometry/ theta, reoh
    common /symmetry/ iamcub
    common /correct/ sig
    common /beaddiabatic/ nbdf1, nbdf2
    common /multiple_ts/ mts
    common /structure/ iamrigid
    prm%oo_eps = oo_eps; prm%oo_sig = oo_sig; prm%oo_gam = oo_gam; prm%rcut = rcut
    prm%alpha = alpha; prm%ecut = ecut; prm%wm = wm; prm%wh = wh
    prm%wrcut = wrcut; prm%walpha = walpha; prm%rkmax = rkmax; prm%kmax = kmax
    prm%iamcub = merge(1, 0, iamcub)
    prm%apot = apot; prm%bpot = bpot; prm%alp = alp; prm%alpb = alpb
    prm%theta = theta; prm%reoh = reoh; prm%sig = sig
    prm%nbdf1 = nbdf1; prm%nbdf2 = nbdf2; prm%mts = max(mts, 1)
    prm%intra_model = 0          ! RPMD_INTRA_AUTO: the reference's own selection rule
    prm%therm = rpmd_therm
    prm%ttau = rpmd_ttau
    prm%rigid = merge(1, 0, iamrigid)
    prm%rat_tol = rpmd_rattle_tol
  end subroutine rpmd_fill_params

  subroutine rpmd_set_thermostat(therm, ttau)
    ! the host's thermostat choice (src/module_thermostat.f90:3-4; the codes of src/main.f90:218-252):
    ! 'PRO' PILE-L, 'PRM' PILE-G, 'PRG' Bussi, 'PRL' local Langevin, 'GLE'; any other code (e.g.
    ! 'AND') leaves the step thermostat-free, as inside the reference's evolve_pi
    character(len=*), intent(in) :: therm
    real(8), intent(in) :: ttau
    select case (therm(1:min(3, len(therm))))
    case ('PRO'); rpmd_therm = 1
    case ('PRM'); rpmd_therm = 2
    case ('PRG'); rpmd_therm = 3
    case ('PRL'); rpmd_therm = 4
    case ('GLE'); rpmd_therm = 5
    case default; rpmd_therm = 0
    end select
    rpmd_ttau = ttau
  end subroutine rpmd_set_thermostat

  subroutine rpmd_set_barostat(baro, taub, patm, dv)
    ! the host's barostat (module barostat, src/module_barostat.f90:2-4, as src/main.f90:257-279 sets
    ! it up): 'BER' or 'MCI', taub, patm and dv in atomic units; resets the MC counters
    character(len=*), intent(in) :: baro
    real(8), intent(in) :: taub, patm, dv
    select case (baro(1:min(3, len(baro))))
    case ('BER'); rpmd_baro = 1
    case ('MCI'); rpmd_baro = 2
    case default
       write(6,*) ' ** Invalid barostat **'
       error stop 'rpmd_hip'
    end select
    rpmd_taub = taub
    rpmd_patm = patm
    rpmd_dv = dv
    if (c_associated(rpmd_ctx)) call rpmd_check(c_set_barostat(rpmd_ctx, int(rpmd_baro, c_int), taub, patm, dv))
  end subroutine rpmd_set_barostat

  subroutine rpmd_barostat_counts(nacc, nmove)
    ! module barostat's nacc and nmove of the MC barostat
    integer, intent(out) :: nacc, nmove
    integer(c_long_long) :: a, m
    a = 0
    m = 0
    if (c_associated(rpmd_ctx)) call rpmd_check(c_barostat_counts(rpmd_ctx, a, m))
    nacc = int(a)
    nmove = int(m)
  end subroutine rpmd_barostat_counts

  subroutine rpmd_replay_uniforms(u)
    ! the MC barostat's ran2 draws of the next steps in the reference's order (the host draws them
    ! with its ran2), instead of the library's stream
    real(8), intent(in) :: u(:)
    if (allocated(rpmd_uniforms)) deallocate(rpmd_uniforms)
    allocate(rpmd_uniforms(size(u)))
    rpmd_uniforms = u
  end subroutine rpmd_replay_uniforms

  subroutine rpmd_replay_noise(noise)
    ! deviates consumed by the next step entry point in the reference's draw order (the host draws
    ! them with its gaussian), instead of the device stream
    real(8), intent(in) :: noise(:)
    if (allocated(rpmd_noise)) deallocate(rpmd_noise)
    allocate(rpmd_noise(size(noise)))
    rpmd_noise = noise
  end subroutine rpmd_replay_noise

  subroutine rpmd_device_noise()
    ! back to the device's own stream
    if (allocated(rpmd_noise)) deallocate(rpmd_noise)
  end subroutine rpmd_device_noise

  subroutine rpmd_noise_arg(np, nn)
    type(c_ptr), intent(out) :: np
    integer(c_long_long), intent(out) :: nn
    if (allocated(rpmd_noise)) then
       np = c_loc(rpmd_noise)
       nn = size(rpmd_noise, kind=c_long_long)
    else
       np = c_null_ptr
       nn = 0
    endif
  end subroutine rpmd_noise_arg

  subroutine rpmd_gle_setup(ns, a, na, nb, dt, beta, gp, seed)
    ! therm_gle_init with the drift matrix given: a(ns+1, ns+1) as the GLE-A file lists it (a(i,:)
    ! = row i), dt the half step the reference passes (src/main.f90:609); gp(3,na,nb,ns+1) the
    ! initial auxiliary momenta (optional: drawn on the device from seed otherwise)
    integer, intent(in) :: ns, na, nb
    real(8), intent(in) :: a(ns + 1, ns + 1), dt, beta
    real(8), intent(in), optional, target :: gp(3, na, nb, ns + 1)
    integer, intent(in), optional :: seed
    real(c_double), target :: at(ns + 1, ns + 1)
    integer(c_long_long) :: sd
    call rpmd_bind(na, nb)
    at = transpose(a)      ! row order in memory, as the C-ABI takes it
    sd = 0
    if (present(seed)) sd = seed
    if (present(gp)) then
       call rpmd_check(c_gle_init(rpmd_ctx, int(ns, c_int), at, int(na, c_int), int(nb, c_int), dt, beta, &
                                  c_loc(gp), sd))
    else
       call rpmd_check(c_gle_init(rpmd_ctx, int(ns, c_int), at, int(na, c_int), int(nb, c_int), dt, beta, &
                                  c_null_ptr, sd))
    endif
  end subroutine rpmd_gle_setup

  subroutine rpmd_therm_gle_init(ttau, na, nb, dt, irun, beta)
    ! therm_gle_init(ttau,na,nb,dt,irun,beta) (src/module_gle.f90:146-227): reads GLE-A from the
    ! working directory as the reference does (ns, then ns+1 rows); C = (nb/beta) I; the auxiliary
    ! momenta are drawn on the device (seeded by irun)
    real(8), intent(in) :: ttau, dt, beta
    integer, intent(in) :: na, nb, irun
    real(8), allocatable :: a(:, :)
    integer :: ns, i, ios
    rpmd_ttau = ttau
    open(121, file='GLE-A', status='old', iostat=ios)
    if (ios .ne. 0) then
       write(0,*) 'COULD NOT FIND GLE-A!'
       error stop 'rpmd_hip'
    endif
    read(121, *) ns
    allocate(a(ns + 1, ns + 1))
    do i = 1, ns + 1
       read(121, *) a(i, :)
    enddo
    close(121)
    call rpmd_gle_setup(ns, a, na, nb, dt, beta, seed=irun)
  end subroutine rpmd_therm_gle_init

  subroutine rpmd_check(code)
    integer(c_int), intent(in) :: code
    character(kind=c_char), pointer :: msg(:)
    character(len=128) :: text
    integer :: i
    if (code .eq. 0) return
    call c_f_pointer(c_strerror(code), msg, [128])
    text = ' '
    do i = 1, 128
       if (msg(i) .eq. c_null_char) exit
       text(i:i) = msg(i)
    enddo
    write(6,*) trim(text)
    error stop 'rpmd_hip'
  end subroutine rpmd_check

  subroutine rpmd_bind(na, nb)
    ! context sized for the largest (na, nb) seen; parameters refreshed every call
    integer, intent(in) :: na, nb
    type(rpmd_params_t) :: prm
    integer(c_int) :: rc
    call rpmd_fill_params(prm)
    if (c_associated(rpmd_ctx) .and. (na .ne. ctx_na .or. nb .gt. ctx_nb)) then
       rc = c_ctx_destroy(rpmd_ctx)
       rpmd_ctx = c_null_ptr
    endif
    if (.not. c_associated(rpmd_ctx)) then
       ctx_na = na
       ctx_nb = max(nb, ctx_nb, 64)
       call rpmd_check(c_ctx_create(prm, int(na, c_int), int(ctx_nb, c_int), int(rpmd_precision, c_int), rpmd_ctx))
       if (rpmd_baro .ne. 0) call rpmd_check(c_set_barostat(rpmd_ctx, int(rpmd_baro, c_int), rpmd_taub, rpmd_patm, &
                                                            rpmd_dv))
    else
       call rpmd_check(c_ctx_set_params(rpmd_ctx, prm))
    endif
  end subroutine rpmd_bind

end module rpmd_hip


subroutine forces(r, v, dvdr, nb, na, boxlxyz, z, virial, iopt)
  use rpmd_hip
  implicit none
  integer nb, na, iopt
  real(8) r(3,na,nb), dvdr(3,na,nb), z(na), boxlxyz(3), v, virial(3,3)
  call rpmd_bind(na, nb)
  call rpmd_check(c_forces(rpmd_ctx, r, int(nb, c_int), int(na, c_int), boxlxyz, z, int(iopt, c_int), &
                           v, dvdr, virial))
end subroutine forces


subroutine full_forces(r, na, nb, v, vew, voo, vint, vir, z, boxlxyz, dvdr, dvdr2)
  use rpmd_hip
  implicit none
  integer na, nb
  real(8) r(3,na,nb), dvdr(3,na,nb), dvdr2(3,na,nb), z(na), boxlxyz(3)
  real(8) v, vew, voo, vint, vir(3,3)
  call rpmd_bind(na, nb)
  call rpmd_check(c_full_forces(rpmd_ctx, r, int(na, c_int), int(nb, c_int), v, vew, voo, vint, vir, z, &
                                boxlxyz, dvdr, dvdr2))
end subroutine full_forces


subroutine realft(data, m, n, mode)
  use rpmd_hip
  implicit none
  integer m, n, mode
  double precision data(m,n)
  ! the force context holds 3*na*nb doubles per array; rebind only if this call does not fit
  if (.not. c_associated(rpmd_ctx) .or. int(m, 8) * n .gt. int(3 * ctx_na, 8) * ctx_nb) &
       call rpmd_bind(3 * ((m + 8) / 9), n)
  call rpmd_check(c_realft(rpmd_ctx, data, int(m, c_int), int(n, c_int), int(mode, c_int)))
end subroutine realft


subroutine evolve_pi(p, r, v, vew, vlj, vint, dvdr, dvdr2, dt, mass, na, nb, boxlxyz, z, beta, vir, vir_lf, irun, nbaro)
  ! one RPMD step (type 'RPMD'); p, r, dvdr, dvdr2 in and out as in the reference.  The thermostat
  ! deviates come from rpmd_replay_noise when set, else from the device stream (irun is not
  ! consumed).  nbaro = 1: the barostat of rpmd_set_barostat acts on the box (in/out) after the
  ! multiple time steps, driven by the incoming vir_lf; its MC draws from rpmd_replay_uniforms
  ! when set, else from the library's stream.
  use rpmd_hip
  implicit none
  integer na, nb, irun, nbaro
  real(8) p(3,na,nb), r(3,na,nb), dvdr(3,na,nb), dvdr2(3,na,nb), mass(na), z(na), boxlxyz(3)
  real(8) v, vew, vlj, vint, dt, beta, vir(3,3), vir_lf(3,3)
  type(c_ptr) :: np, up
  integer(c_long_long) :: nn
  integer(c_int) :: nu
  call rpmd_bind(na, nb)
  call rpmd_noise_arg(np, nn)
  if (nbaro .ne. 0) then
     up = c_null_ptr
     nu = 0
     if (allocated(rpmd_uniforms)) then
        up = c_loc(rpmd_uniforms)
        nu = size(rpmd_uniforms, kind=c_int)
     endif
     call rpmd_check(c_evolve_pi_npt(rpmd_ctx, p, r, v, vew, vlj, vint, dvdr, dvdr2, dt, mass, int(na, c_int), &
                                     int(nb, c_int), boxlxyz, z, beta, vir, vir_lf, np, nn, up, nu))
     return
  endif
  call rpmd_check(c_evolve_pi(rpmd_ctx, p, r, v, vew, vlj, vint, dvdr, dvdr2, dt, mass, int(na, c_int), &
                              int(nb, c_int), boxlxyz, z, beta, vir, vir_lf, np, nn))
end subroutine evolve_pi


subroutine evolve_rig_pi(p, r, v, v_lf, v_hf, dvdr, dvdr2, dt, mass, na, nb, boxlxyz, z, beta, vir, vir_lf, irun, &
                         nbaro)
  ! one rigid RPMD step (/structure/ iamrigid); vir in and out as in the reference; nbaro = 1: the
  ! barostat of rpmd_set_barostat after the step (box in/out)
  use rpmd_hip
  implicit none
  integer na, nb, irun, nbaro
  real(8) p(3,na,nb), r(3,na,nb), dvdr(3,na,nb), dvdr2(3,na,nb), mass(na), z(na), boxlxyz(3)
  real(8) v, v_lf, v_hf, dt, beta, vir(3,3), vir_lf(3,3)
  type(c_ptr) :: np, up
  integer(c_long_long) :: nn
  integer(c_int) :: nu
  call rpmd_bind(na, nb)
  call rpmd_noise_arg(np, nn)
  if (nbaro .ne. 0) then
     up = c_null_ptr
     nu = 0
     if (allocated(rpmd_uniforms)) then
        up = c_loc(rpmd_uniforms)
        nu = size(rpmd_uniforms, kind=c_int)
     endif
     call rpmd_check(c_evolve_rig_pi_npt(rpmd_ctx, p, r, v, v_lf, v_hf, dvdr, dvdr2, dt, mass, int(na, c_int), &
                                         int(nb, c_int), boxlxyz, z, beta, vir, vir_lf, np, nn, up, nu))
     return
  endif
  call rpmd_check(c_evolve_rig_pi(rpmd_ctx, p, r, v, v_lf, v_hf, dvdr, dvdr2, dt, mass, int(na, c_int), &
                                  int(nb, c_int), boxlxyz, z, beta, vir, vir_lf, np, nn))
end subroutine evolve_rig_pi


subroutine run_driver(p, r, dvdr, dvdr2, na, nb, boxlxyz, z, beta, dt, mass, irun, itst, pt, pb, print)
  ! the i-PI client: reads the server address from the file serverdata, as the reference does, and
  ! serves the server with full_forces on one bead until it disconnects; p, r, dvdr, dvdr2, beta,
  ! dt, mass, irun, itst, pt, pb and print are accepted for the reference's argument list
  use rpmd_hip
  implicit none
  integer na, nb, irun, itst, pt, pb, print
  real(8) p(3,na,nb), r(3,na,nb), dvdr(3,na,nb), dvdr2(3,na,nb), z(na), boxlxyz(3), beta, dt, mass(na)
  character(1024) :: serveraddr
  integer(c_int) :: served
  if (nb .ne. 1) then
     write(*,*) "i-PI mode must have nb = 1!"
     stop
  endif
  write(*,*) "@ i-PI DRIVER BEING LOADED"
  open(12, file="serverdata")
  read(12, *) serveraddr
  close(12)
  write(*,*) "SERVERDATA ", trim(serveraddr)
  call rpmd_bind(na, nb)
  call rpmd_check(c_ipi_run(rpmd_ctx, trim(serveraddr)//c_null_char, z, int(na, c_int), served))
end subroutine run_driver
