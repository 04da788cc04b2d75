! rpmd_hip.f90 -- ISO_C_BINDING layer between the reference's Fortran host and librpmd_mi355x.so.
!
! Link this file in place of the reference's md_forces.f90 (forces, full_forces) and fourier.f90
! (realft): the external subroutines below keep the reference's names and dummy-argument lists
!   forces(r,v,dvdr,nb,na,boxlxyz,z,virial,iopt)                src/md_forces.f90:1
!   full_forces(r,na,nb,v,vew,voo,vint,vir,z,boxlxyz,dvdr,dvdr2) src/md_forces.f90:305
!   realft(data,m,n,mode)                                       src/fourier.f90:16
! so every existing call site compiles and links unchanged.  Model parameters are read from the
! same COMMON blocks the reference routines read (src/main.f90:53-72) on every call; a context
! for the largest (na, nb) seen is created on first use.  A non-zero status from the library
! stops the program with the reference's message, as the reference's STOP statements do.
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
  end interface

  type(c_ptr), save :: rpmd_ctx = c_null_ptr
  integer, save :: ctx_na = 0, ctx_nb = 0
  ! RPMD_PREC_FP64 = 0, RPMD_PREC_MIXED = 1; set before the first call to choose the mode
  integer, save :: rpmd_precision = 0

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
    prm%therm = 0
    prm%ttau = 0.d0
    prm%rigid = merge(1, 0, iamrigid)
  end subroutine rpmd_fill_params

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
