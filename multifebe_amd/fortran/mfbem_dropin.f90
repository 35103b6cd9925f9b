!! Drop-in replacements of four MultiFEBE routines on libmfbem (ISO_C_BINDING over include/mfbem.h).
!!
!! Compiled into the reference program in place of its own files, these keep the reference's
!! argument lists, so its driver calls them unchanged:
!!   solve_lse_r / solve_lse_c                   (src/solve_lse_r.f90:25-46, src/solve_lse_c.f90:25-46)
!!     -> mfb_solve_lse_r / mfb_solve_lse_c: the LU and solve on the GPU; the factor stays on the
!!        device for factorize = .false. re-solves (src/multifebe.f90:119);
!!   assemble_bem_staela_equation / assemble_bem_harela_equation
!!     (src/assemble_bem_staela_equation.f90:22-48, src/assemble_bem_harela_equation.f90:22-53)
!!     -> mfb_scatter_bem_*_equation with the reference's own DOF tables (node%row, node%col).
!! Logicals travel as int, character(1) as one c_char; the condition and refine stages run in the
!! library (?gecon / ?gerfs on the GPU factors) and their messages are printed here.  A nonzero return prints the library's
!! message and stops, as the reference's fbem_error_message does.  MFBEM_PRECISION in the
!! environment selects the solve's mode: 0 (default) FP64 LU, 1 FP32 LU + FP64 refinement, 2 FP32
!! LU with split FP16 trailing updates + FP64 refinement.

module mfbem_capi
  use iso_c_binding
  implicit none
  type, bind(c) :: mfb_solve_report
    integer(c_int) :: info, precision_mode, refinement_steps, fell_back
    real(c_double) :: backward_error, factor_ms, solve_ms, rcond, ferr
  end type
  interface
    function mfb_last_error() bind(c, name='mfb_last_error')
      import :: c_ptr
      type(c_ptr) :: mfb_last_error
    end function
    integer(c_int) function mfb_solve_lse_r(n, A, ipiv, aodim, Ao, equed, r, c, nrhs, b, fact, scal, &
                                            cond, refi, prec, rep) bind(c, name='mfb_solve_lse_r')
      import :: c_int, c_double, c_char, mfb_solve_report
      integer(c_int), value :: n, aodim, nrhs, fact, scal, cond, refi, prec
      real(c_double) :: A(*), Ao(*), r(*), c(*), b(*)
      integer(c_int) :: ipiv(*)
      character(kind=c_char) :: equed(*)
      type(mfb_solve_report) :: rep
    end function
    integer(c_int) function mfb_solve_lse_c(n, A, ipiv, aodim, Ao, equed, r, c, nrhs, b, fact, scal, &
                                            cond, refi, prec, rep) bind(c, name='mfb_solve_lse_c')
      import :: c_int, c_double, c_char, mfb_solve_report
      integer(c_int), value :: n, aodim, nrhs, fact, scal, cond, refi, prec
      complex(c_double) :: A(*), Ao(*), b(*)
      real(c_double) :: r(*), c(*)
      integer(c_int) :: ipiv(*)
      character(kind=c_char) :: equed(*)
      type(mfb_solve_report) :: rep
    end function
    integer(c_int) function mfb_scatter_bem_staela_equation(ne, rows, cols, ctype, cvalue, n_fn, rev, &
                                                            h, g, A, lda, b) &
                                                            bind(c, name='mfb_scatter_bem_staela_equation')
      import :: c_int, c_double, c_long_long
      integer(c_int), value :: ne, rev
      integer(c_int) :: rows(*), cols(*), ctype(*)
      real(c_double) :: cvalue(*), n_fn(*), h(*), g(*), A(*), b(*)
      integer(c_long_long), value :: lda
    end function
    integer(c_int) function mfb_scatter_bem_harela_equation(ne, rows, cols, ctype, cvalue, n_fn, rev, &
                                                            h, g, u_inc, t_inc, A, lda, b) &
                                                            bind(c, name='mfb_scatter_bem_harela_equation')
      import :: c_int, c_double, c_long_long, c_ptr
      integer(c_int), value :: ne, rev
      integer(c_int) :: rows(*), cols(*), ctype(*)
      complex(c_double) :: cvalue(*), h(*), g(*), A(*), b(*)
      real(c_double) :: n_fn(*)
      type(c_ptr), value :: u_inc, t_inc
      integer(c_long_long), value :: lda
    end function
  end interface
contains
  !! stop with the library's message (the reference prints its message and stops)
  subroutine mfbem_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    if (rc == 0) return
    call c_f_pointer(mfb_last_error(), msg, [1024])
    k = 0
    do while (k < 1024)
      if (msg(k + 1) == c_null_char) exit
      k = k + 1
    end do
    write(0, '(a,a,i0,a,1024a)') what, ' failed (code ', rc, '): ', msg(1:k)
    error stop 1
  end subroutine
  integer(c_int) function mfbem_precision()
    character(len=16) :: v
    integer :: st
    mfbem_precision = 0
    call get_environment_variable('MFBEM_PRECISION', v, status=st)
    if (st == 0) then
      if (trim(v) == '1') mfbem_precision = 1
      if (trim(v) == '2') mfbem_precision = 2
    end if
  end function
  !! the messages of the reference's condition and refine stages (src/solve_lse_c.f90:158-161,
  !! 194-199) at verbose level 2; the forward and backward errors are the largest over the
  !! right-hand sides
  subroutine mfbem_stage_messages(rep, condition, refine, n_rhs)
    use iso_fortran_env, only: output_unit
    use fbem_string_handling, only: verbose_level, fbem_timestamp_message
    type(mfb_solve_report), intent(in) :: rep
    logical, intent(in) :: condition, refine
    integer, intent(in) :: n_rhs
    integer :: i
    if (verbose_level < 2) return
    if (condition .and. rep%rcond > 0.0d0) then
      call fbem_timestamp_message(output_unit, 2)
      write(output_unit, '(a17,e25.16)') 'Condition number:', 1.0d0/rep%rcond
    end if
    if (refine) then
      do i = 1, n_rhs
        call fbem_timestamp_message(output_unit, 2)
        write(output_unit, '(a30,i11,a5,e25.16,a5,e25.16)') 'Solution refinement done: rhs=', i, &
              ', berr=', rep%backward_error, ', ferr=', rep%ferr
      end do
    end if
  end subroutine
  integer(c_int) function lgc(l)
    logical, intent(in) :: l
    lgc = 0
    if (l) lgc = 1
  end function
end module mfbem_capi

subroutine solve_lse_r(n_dof,A,ipiv,Aodim,Ao,equed,r,c,n_rhs,b,factorize,scaling,condition,refine)
  use iso_fortran_env
  use iso_c_binding
  use mfbem_capi
  implicit none
  integer           :: n_dof
  real(kind=real64) :: A(n_dof,n_dof)
  integer           :: ipiv(n_dof)
  integer           :: Aodim
  real(kind=real64) :: Ao(Aodim,Aodim)
  character(len=1)  :: equed
  real(kind=real64) :: r(n_dof)
  real(kind=real64) :: c(n_dof)
  integer           :: n_rhs
  real(kind=real64) :: b(n_dof,n_rhs)
  logical           :: factorize, scaling, condition, refine
  character(kind=c_char) :: eq(2)
  type(mfb_solve_report) :: rep
  eq(1) = equed
  eq(2) = c_null_char
  call mfbem_check(mfb_solve_lse_r(n_dof, A, ipiv, Aodim, Ao, eq, r, c, n_rhs, b, lgc(factorize), &
                                   lgc(scaling), lgc(condition), lgc(refine), mfbem_precision(), rep), &
                   'solve_lse_r')
  equed = eq(1)
  call mfbem_stage_messages(rep, condition .and. factorize, refine, n_rhs)
end subroutine solve_lse_r

subroutine solve_lse_c(n_dof,A,ipiv,Aodim,Ao,equed,r,c,n_rhs,b,factorize,scaling,condition,refine)
  use iso_fortran_env
  use iso_c_binding
  use mfbem_capi
  implicit none
  integer              :: n_dof
  complex(kind=real64) :: A(n_dof,n_dof)
  integer              :: ipiv(n_dof)
  integer              :: Aodim
  complex(kind=real64) :: Ao(Aodim,Aodim)
  character(len=1)     :: equed
  real(kind=real64)    :: r(n_dof)
  real(kind=real64)    :: c(n_dof)
  integer              :: n_rhs
  complex(kind=real64) :: b(n_dof,n_rhs)
  logical              :: factorize, scaling, condition, refine
  character(kind=c_char) :: eq(2)
  type(mfb_solve_report) :: rep
  eq(1) = equed
  eq(2) = c_null_char
  call mfbem_check(mfb_solve_lse_c(n_dof, A, ipiv, Aodim, Ao, eq, r, c, n_rhs, b, lgc(factorize), &
                                   lgc(scaling), lgc(condition), lgc(refine), mfbem_precision(), rep), &
                   'solve_lse_c')
  equed = eq(1)
  call mfbem_stage_messages(rep, condition .and. factorize, refine, n_rhs)
end subroutine solve_lse_c

!! The reference's module globals (node, element, boundary, A_r / A_c, b_r / b_c) are read here and
!! handed to the library as plain tables; BE-ordinary boundaries only (the library's scope).
subroutine assemble_bem_staela_equation(kr,sb_int,sb_int_reversion,se_int,se_int_n_nodes,sn_col,eq_index,hp,gp,hm,gm)
  use iso_fortran_env
  use iso_c_binding
  use fbem_data_structures
  use problem_variables
  use mfbem_capi
  implicit none
  integer           :: kr
  integer           :: sb_int
  logical           :: sb_int_reversion
  integer           :: se_int
  integer           :: se_int_n_nodes
  integer           :: sn_col
  integer           :: eq_index
  real(kind=real64) :: hp(se_int_n_nodes,problem%n,problem%n)
  real(kind=real64) :: gp(se_int_n_nodes,problem%n,problem%n)
  real(kind=real64) :: hm(se_int_n_nodes,problem%n,problem%n)
  real(kind=real64) :: gm(se_int_n_nodes,problem%n,problem%n)
  integer(c_int) :: rows(3), cols(6*se_int_n_nodes), ctype(3*se_int_n_nodes)
  real(c_double) :: cv(3*se_int_n_nodes), nfn(3*se_int_n_nodes)
  integer :: kn, sn, k
  if ((boundary(sb_int)%coupling.ne.fbem_boundary_coupling_be).or.&
      (boundary(sb_int)%class.ne.fbem_boundary_class_ordinary).or.(problem%n.ne.3)) then
    write(0,*) 'assemble_bem_staela_equation: libmfbem scatters ordinary 3D BE boundaries only'
    error stop 1
  end if
  do k=1,3
    rows(k)=node(sn_col)%row(k,eq_index)-1
  end do
  do kn=1,se_int_n_nodes
    sn=element(se_int)%node(kn)
    do k=1,6
      cols(6*(kn-1)+k)=node(sn)%col(k,1)-1
    end do
    do k=1,3
      ctype(3*(kn-1)+k)=node(sn)%ctype(k,1)
      cv(3*(kn-1)+k)=node(sn)%cvalue_r(k,1,1)
      nfn(3*(kn-1)+k)=0.d0
      if (allocated(node(sn)%n_fn)) nfn(3*(kn-1)+k)=node(sn)%n_fn(k)
    end do
  end do
  call mfbem_check(mfb_scatter_bem_staela_equation(se_int_n_nodes, rows, cols, ctype, cv, nfn, &
                                                   lgc(sb_int_reversion), hp, gp, A_r, &
                                                   int(size(A_r,1),c_long_long), b_r), &
                   'assemble_bem_staela_equation')
end subroutine assemble_bem_staela_equation

subroutine assemble_bem_harela_equation(omega,kr,sb_int,sb_int_reversion,se_int,se_int_n_nodes,sn_col,eq_index,hp,gp,hm,gm,u_inc,t_inc)
  use iso_fortran_env
  use iso_c_binding
  use fbem_data_structures
  use problem_variables
  use mfbem_capi
  implicit none
  real(kind=real64)    :: omega
  integer              :: kr
  integer              :: sb_int
  logical              :: sb_int_reversion
  integer              :: se_int
  integer              :: se_int_n_nodes
  integer              :: sn_col
  integer              :: eq_index
  complex(kind=real64) :: hp(se_int_n_nodes,problem%n,problem%n)
  complex(kind=real64) :: gp(se_int_n_nodes,problem%n,problem%n)
  complex(kind=real64) :: hm(se_int_n_nodes,problem%n,problem%n)
  complex(kind=real64) :: gm(se_int_n_nodes,problem%n,problem%n)
  complex(kind=real64), target :: u_inc(problem%n,se_int_n_nodes)
  complex(kind=real64), target :: t_inc(problem%n,se_int_n_nodes)
  integer(c_int) :: rows(3), cols(6*se_int_n_nodes), ctype(3*se_int_n_nodes)
  complex(c_double) :: cv(3*se_int_n_nodes)
  real(c_double) :: nfn(3*se_int_n_nodes)
  integer :: kn, sn, k
  if ((boundary(sb_int)%coupling.ne.fbem_boundary_coupling_be).or.&
      (boundary(sb_int)%class.ne.fbem_boundary_class_ordinary).or.(problem%n.ne.3)) then
    write(0,*) 'assemble_bem_harela_equation: libmfbem scatters ordinary 3D BE boundaries only'
    error stop 1
  end if
  do k=1,3
    rows(k)=node(sn_col)%row(k,eq_index)-1
  end do
  do kn=1,se_int_n_nodes
    sn=element(se_int)%node(kn)
    do k=1,6
      cols(6*(kn-1)+k)=node(sn)%col(k,1)-1
    end do
    do k=1,3
      ctype(3*(kn-1)+k)=node(sn)%ctype(k,1)
      cv(3*(kn-1)+k)=node(sn)%cvalue_c(k,1,1)
      nfn(3*(kn-1)+k)=0.d0
      if (allocated(node(sn)%n_fn)) nfn(3*(kn-1)+k)=node(sn)%n_fn(k)
    end do
  end do
  if (region(kr)%n_incidentfields.gt.0) then
    call mfbem_check(mfb_scatter_bem_harela_equation(se_int_n_nodes, rows, cols, ctype, cv, nfn, &
                                                     lgc(sb_int_reversion), hp, gp, c_loc(u_inc), &
                                                     c_loc(t_inc), A_c, int(size(A_c,1),c_long_long), &
                                                     b_c), 'assemble_bem_harela_equation')
  else
    call mfbem_check(mfb_scatter_bem_harela_equation(se_int_n_nodes, rows, cols, ctype, cv, nfn, &
                                                     lgc(sb_int_reversion), hp, gp, c_null_ptr, &
                                                     c_null_ptr, A_c, int(size(A_c,1),c_long_long), &
                                                     b_c), 'assemble_bem_harela_equation')
  end if
end subroutine assemble_bem_harela_equation
