!! Golden-vector harness around the reference FBEM library (test infrastructure only).
!!
!! This program is the project's own code.  It links the reference's library modules
!! (lib/fbem/src/*.f90, compiled unmodified by oracle/Makefile) and calls them on inputs read
!! from stdin, printing every result with 17 significant digits so that tests/golden/make_golden.py
!! can store them as fixtures.  It is never shipped to a GPU host and never used by the product.
!!
!! Modes (first line of stdin):
!!   gl            Gauss-Legendre [-1,1] and [0,1] tables (quad_rules.f90:79-136)
!!   zexp          fbem_zexp_decomposed (numerical.f90:1258-1330)
!!   freeterm      fbem_bem_staela3d_sbie_freeterm (bem_staela3d.f90:228-405)
!!   elem          element data as src/build_data_of_be_elements.f90:70-112 computes it
!!   pair_sta      decision trace + fbem_bem_staela3d_sbie_auto (bem_staela3d.f90:524-589)
!!   pair_har      decision trace + fbem_bem_harela3d_sbie_auto (bem_harela3d.f90:1474-1538)
!!   assemble_sta  whole-boundary A,b for a nodal-collocation quad9 mesh, following the loop order of
!!                 src/build_lse_mechanics_bem_staela.f90:112-549,659-1034 and the BE-ordinary scatter
!!                 rule of src/assemble_bem_staela_equation.f90:69-98
!!   assemble_har  same for src/build_lse_mechanics_bem_harela.f90 + assemble_bem_harela_equation.f90:78-113
program ref_harness
  use iso_fortran_env
  use fbem_numerical
  use fbem_shape_functions
  use fbem_quad_rules
  use fbem_quasisingular_integration
  use fbem_geometry
  use fbem_bem_general
  use fbem_bem_staela3d
  use fbem_bem_harela3d
  implicit none
  character(len=32) :: mode
  call fbem_init_shape_functions_module
  read(*,*) mode
  select case (trim(mode))
    case ('gl')
      call do_gl()
    case ('zexp')
      call do_zexp()
    case ('freeterm')
      call do_freeterm()
    case ('elem')
      call do_elem()
    case ('pair_sta')
      call do_pair(.false.)
    case ('pair_har')
      call do_pair(.true.)
    case ('assemble_sta')
      call do_assemble(.false.)
    case ('assemble_har')
      call do_assemble(.true.)
    case default
      stop 'unknown mode'
  end select

contains

  subroutine pr(v)
    real(kind=real64) :: v
    write(*,'(es26.17e3)') v
  end subroutine pr

  subroutine pri(i)
    integer :: i
    write(*,'(i0)') i
  end subroutine pri

  subroutine do_gl()
    integer :: n, k
    do n=1,32
      do k=1,n
        write(*,'(4es26.17e3)') gl11_xi(k,n), gl11_w(k,n), gl01_xi(k,n), gl01_w(k,n)
      end do
    end do
  end subroutine do_gl

  subroutine do_zexp()
    integer :: nz, k, j
    real(kind=real64) :: zr, zi
    complex(kind=real64) :: z(1), E(0:6,1)
    read(*,*) nz
    do k=1,nz
      read(*,*) zr, zi
      z(1)=cmplx(zr,zi,kind=real64)
      call fbem_zexp_decomposed(1,z,E)
      do j=0,6
        write(*,'(2es26.17e3)') dble(E(j,1)), dimag(E(j,1))
      end do
    end do
  end subroutine do_zexp

  subroutine do_freeterm()
    integer :: m, k, l
    real(kind=real64), allocatable :: n(:,:), t(:,:)
    real(kind=real64) :: tol, nu, c(3,3)
    read(*,*) m, tol, nu
    allocate (n(3,m), t(3,m))
    do k=1,m
      read(*,*) n(:,k)
      read(*,*) t(:,k)
    end do
    call fbem_bem_staela3d_sbie_freeterm(m,n,t,tol,nu,c)
    do l=1,3
      do k=1,3
        call pr(c(k,l))
      end do
    end do
  end subroutine do_freeterm

  !! Element data exactly as build_data_of_be_elements.f90:70-112 derives it.
  subroutine element_data(x, qsi_re, csize, n_phi, centre, radius)
    real(kind=real64) :: x(3,9), qsi_re, csize, centre(3), radius
    integer :: n_phi
    logical :: ok
    csize=fbem_characteristic_length(3,fbem_quad9,x,1.d-9)
    call fbem_qs_phijac_ngp_2d(3,fbem_quad9,fbem_quad9,x,qsi_re,n_phi,ok)
    call fbem_geometry_element_ball(3,fbem_quad9,x,n_phi,centre,radius)
  end subroutine element_data

  subroutine do_elem()
    real(kind=real64) :: x(3,9), qsi_re, csize, centre(3), radius
    integer :: n_phi, k
    read(*,*) qsi_re
    do k=1,9
      read(*,*) x(:,k)
    end do
    call element_data(x, qsi_re, csize, n_phi, centre, radius)
    call pr(csize)
    call pri(n_phi)
    call pr(centre(1)); call pr(centre(2)); call pr(centre(3))
    call pr(radius)
  end subroutine do_elem

  subroutine setup_element(e, x, qsi_re, n_ps, ps_gln)
    type(fbem_bem_element) :: e
    real(kind=real64) :: x(3,9), qsi_re
    integer :: n_ps, ps_gln(n_ps)
    real(kind=real64) :: csize, centre(3), radius
    integer :: n_phi
    call element_data(x, qsi_re, csize, n_phi, centre, radius)
    call e%init
    e%gtype=fbem_quad9
    e%d=2
    e%n_gnodes=9
    e%n=3
    allocate (e%x(3,9))
    e%x=x
    e%ptype=fbem_quad9
    e%ptype_delta=0.d0
    e%n_pnodes=9
    e%stype=fbem_quad9
    e%stype_delta=0.d0
    e%n_snodes=9
    e%cl=csize
    e%gln_far=n_phi
    allocate (e%bball_centre(3))
    e%bball_centre=centre
    e%bball_radius=radius
    call e%init_precalculated_datasets(n_ps,ps_gln)
  end subroutine setup_element

  !! The decision sequence of fbem_bem_staela3d_sbie_auto (bem_staela3d.f90:549-587), replayed with the
  !! same library calls so that the chosen path can be compared as well as the values.
  subroutine trace_decision(e, x_i, qsp, delta, barxi, d, gln_near, gln, ps)
    type(fbem_bem_element) :: e
    real(kind=real64) :: x_i(3), barxi(2), d, r(3), rmin
    type(fbem_qs_parameters) :: qsp
    integer :: delta, gln_near, gln, ps, method, i
    r=e%bball_centre-x_i
    rmin=sqrt(dot_product(r,r))-e%bball_radius
    gln_near=-1
    gln=-1
    ps=-1
    if (rmin.gt.(4.d0*e%bball_radius)) then
      delta=0
      barxi=0.d0
      d=rmin/e%cl
    else
      call fbem_nearest_element_point_bem(3,e%gtype,e%x,e%cl,x_i,barxi,rmin,d,method)
      if (d.le.1.d-12) then
        delta=1
      else
        delta=0
      end if
    end if
    if (delta.eq.0) then
      gln_near=fbem_qs_n_estimation_standard(e%n,e%gtype,5,qsp,d,barxi)
      gln=max(e%gln_far,gln_near)
      ps=0
      if ((gln.le.e%ps_gln_max).and.(gln_near.gt.0)) then
        do i=1,e%n_ps
          if (e%ps_gln(i).ge.gln) then
            ps=i
            exit
          end if
        end do
      end if
    end if
  end subroutine trace_decision

  subroutine do_pair(harmonic)
    logical :: harmonic
    real(kind=real64) :: x(3,9), x_i(3), mu_r, nu_r, qsi_re, barxi(2), d
    real(kind=real64) :: lr, li, mr, mi, rho, omega
    integer :: ns, irev, n_ps, k, l, m, delta, gln_near, gln, ps
    integer, allocatable :: ps_gln(:)
    logical :: rev
    type(fbem_bem_element) :: e
    type(fbem_qs_parameters) :: qsp
    type(fbem_bem_harela3d_parameters) :: p
    real(kind=real64) :: h(9,3,3), g(9,3,3)
    complex(kind=real64) :: hc(9,3,3), gc(9,3,3)
    read(*,*) qsi_re, ns, irev
    read(*,*) n_ps
    allocate (ps_gln(n_ps))
    read(*,*) ps_gln
    if (harmonic) then
      read(*,*) lr, li, mr, mi, rho, omega
    else
      read(*,*) mu_r, nu_r
    end if
    do k=1,9
      read(*,*) x(:,k)
    end do
    read(*,*) x_i
    rev=(irev.ne.0)
    call fbem_qs_calculate_parameters(qsi_re,qsp)
    call setup_element(e, x, qsi_re, n_ps, ps_gln)
    call trace_decision(e, x_i, qsp, delta, barxi, d, gln_near, gln, ps)
    call pri(delta); call pri(gln_near); call pri(gln); call pri(ps)
    call pr(barxi(1)); call pr(barxi(2)); call pr(d)
    if (harmonic) then
      call fbem_bem_harela3d_calculate_parameters(cmplx(lr,li,kind=real64),cmplx(mr,mi,kind=real64),rho,omega,p)
      call fbem_bem_harela3d_sbie_auto(e,rev,x_i,p,qsp,ns,hc,gc)
      do m=1,3
        do l=1,3
          do k=1,9
            write(*,'(4es26.17e3)') dble(hc(k,l,m)), dimag(hc(k,l,m)), dble(gc(k,l,m)), dimag(gc(k,l,m))
          end do
        end do
      end do
    else
      call fbem_bem_staela3d_sbie_auto(e,rev,x_i,mu_r,nu_r,qsp,ns,h,g)
      do m=1,3
        do l=1,3
          do k=1,9
            write(*,'(2es26.17e3)') h(k,l,m), g(k,l,m)
          end do
        end do
      end do
    end if
  end subroutine do_pair

  !! Unit boundary tangent "tbp" of a quad9 at local node kn (geometry.f90:657-870 for quads).
  function quad9_tbp(x, kn) result(t)
    real(kind=real64) :: x(3,9), t(3)
    integer :: kn
    real(kind=real64) :: xi(2), t1(3), t2(3)
    xi=quad9_xi_node(kn)
    t1=fbem_utangent_xi1(3,fbem_quad9,x,xi)
    t2=fbem_utangent_xi2(3,fbem_quad9,x,xi)
    select case (kn)
      case (1,5)
        t=t1
      case (2,6)
        t=t2
      case (3,7)
        t=-t1
      case (4,8)
        t=-t2
      case default
        t=0.d0
    end select
  end function quad9_tbp

  function quad9_xi_node(kn) result(xi)
    integer :: kn
    real(kind=real64) :: xi(2)
    real(kind=real64), parameter :: xn(2,9)=reshape((/-1.d0,-1.d0, 1.d0,-1.d0, 1.d0,1.d0, -1.d0,1.d0, &
                                                      0.d0,-1.d0, 1.d0,0.d0, 0.d0,1.d0, -1.d0,0.d0, 0.d0,0.d0/),(/2,9/))
    xi=xn(:,kn)
  end function quad9_xi_node

  subroutine do_assemble(harmonic)
    logical :: harmonic
    integer :: nn, ne, n_ps, ns, k, ke, kn, kn2, ie, jn, il, ik, row, col, m, kc
    integer, allocatable :: conn(:,:), ctype(:,:), ps_gln(:), node_first_elem(:), node_first_loc(:)
    real(kind=real64), allocatable :: xn(:,:), cval(:,:), nfn(:,:), A(:,:), b(:), nset(:,:), tset(:,:)
    complex(kind=real64), allocatable :: Ac(:,:), bc(:)
    real(kind=real64) :: mu_r, nu_r, qsi_re, lr, li, mr, mi, rho, omega, x(3,9), x_i(3), xi(2), cfree(3,3), tol
    real(kind=real64) :: h(9,3,3), g(9,3,3), nv(3)
    complex(kind=real64) :: hc(9,3,3), gc(9,3,3), cfree_c(3,3), nu_c
    logical, allocatable :: collocated(:)
    type(fbem_bem_element) :: e
    type(fbem_qs_parameters) :: qsp
    type(fbem_bem_harela3d_parameters) :: p
    read(*,*) qsi_re, ns, tol
    read(*,*) n_ps
    allocate (ps_gln(n_ps))
    read(*,*) ps_gln
    if (harmonic) then
      read(*,*) lr, li, mr, mi, rho, omega, nu_r
    else
      read(*,*) mu_r, nu_r
    end if
    read(*,*) nn, ne
    allocate (xn(3,nn), conn(9,ne), ctype(3,nn), cval(3,nn), nfn(3,nn), collocated(nn))
    allocate (node_first_elem(nn), node_first_loc(nn))
    do k=1,nn
      read(*,*) xn(:,k)
    end do
    do k=1,ne
      read(*,*) conn(:,k)
    end do
    do k=1,nn
      read(*,*) ctype(:,k)
      read(*,*) cval(:,k)
    end do
    call fbem_qs_calculate_parameters(qsi_re,qsp)
    ! Node normals: normalised sum of the element normals at the node (build_data_at_functional_nodes.f90:346-367)
    nfn=0.d0
    node_first_elem=0
    do ke=1,ne
      do kn=1,9
        x=xn(:,conn(:,ke))
        nv=fbem_unormal3d(fbem_quad9,x,quad9_xi_node(kn))
        nfn(:,conn(kn,ke))=nfn(:,conn(kn,ke))+nv
        if (node_first_elem(conn(kn,ke)).eq.0) then
          node_first_elem(conn(kn,ke))=ke
          node_first_loc(conn(kn,ke))=kn
        end if
      end do
    end do
    do k=1,nn
      nfn(:,k)=nfn(:,k)/sqrt(dot_product(nfn(:,k),nfn(:,k)))
    end do
    if (harmonic) then
      allocate (Ac(3*nn,3*nn), bc(3*nn))
      Ac=(0.d0,0.d0)
      bc=(0.d0,0.d0)
      call fbem_bem_harela3d_calculate_parameters(cmplx(lr,li,kind=real64),cmplx(mr,mi,kind=real64),rho,omega,p)
    else
      allocate (A(3*nn,3*nn), b(3*nn))
      A=0.d0
      b=0.d0
    end if
    ! Pair loop: integration element outer, collocation nodes inner (each node once per element)
    do ke=1,ne
      x=xn(:,conn(:,ke))
      if (allocated(e%x)) deallocate(e%x)
      if (allocated(e%bball_centre)) deallocate(e%bball_centre)
      call setup_element(e, x, qsi_re, n_ps, ps_gln)
      collocated=.false.
      do ie=1,ne
        do kn=1,9
          jn=conn(kn,ie)
          if (collocated(jn)) cycle
          collocated(jn)=.true.
          x_i=xn(:,jn)
          if (harmonic) then
            call fbem_bem_harela3d_sbie_auto(e,.false.,x_i,p,qsp,ns,hc,gc)
            call scatter_c(jn, conn(:,ke), hc, gc)
          else
            call fbem_bem_staela3d_sbie_auto(e,.false.,x_i,mu_r,nu_r,qsp,ns,h,g)
            call scatter_r(jn, conn(:,ke), h, g)
          end if
        end do
      end do
    end do
    ! Free terms (build_lse_mechanics_bem_staela.f90:152-549): first element holding the node
    collocated=.false.
    do ke=1,ne
      do kn=1,9
        jn=conn(kn,ke)
        if (collocated(jn)) cycle
        collocated(jn)=.true.
        xi=quad9_xi_node(kn)
        if (kn.le.8) then
          m=0
          do ie=1,ne
            do kn2=1,9
              if (conn(kn2,ie).eq.jn) m=m+1
            end do
          end do
          allocate (nset(3,m), tset(3,m))
          m=0
          do ie=1,ne
            do kn2=1,9
              if (conn(kn2,ie).eq.jn) then
                m=m+1
                x=xn(:,conn(:,ie))
                nset(:,m)=fbem_unormal3d(fbem_quad9,x,quad9_xi_node(kn2))
                tset(:,m)=quad9_tbp(x,kn2)
              end if
            end do
          end do
          if (harmonic) then
            nu_c=cmplx(nu_r,0.d0,kind=real64)
            call fbem_bem_harela3d_sbie_freeterm(m,nset,tset,tol,nu_c,cfree_c)
          else
            call fbem_bem_staela3d_sbie_freeterm(m,nset,tset,tol,nu_r,cfree)
          end if
          deallocate (nset, tset)
        else
          cfree=0.d0
          do kc=1,3
            cfree(kc,kc)=0.5d0
          end do
          cfree_c=cmplx(cfree,0.d0,kind=real64)
        end if
        if (harmonic) then
          hc=(0.d0,0.d0)
          gc=(0.d0,0.d0)
          hc(kn,:,:)=cfree_c
          call scatter_c(jn, conn(:,ke), hc, gc)
        else
          h=0.d0
          g=0.d0
          h(kn,:,:)=cfree
          call scatter_r(jn, conn(:,ke), h, g)
        end if
      end do
    end do
    if (harmonic) then
      do col=1,3*nn
        do row=1,3*nn
          write(*,'(2es26.17e3)') dble(Ac(row,col)), dimag(Ac(row,col))
        end do
      end do
      do row=1,3*nn
        write(*,'(2es26.17e3)') dble(bc(row)), dimag(bc(row))
      end do
    else
      do col=1,3*nn
        do row=1,3*nn
          call pr(A(row,col))
        end do
      end do
      do row=1,3*nn
        call pr(b(row))
      end do
    end if

  contains

    !! BE ordinary scatter (assemble_bem_staela_equation.f90:69-98): one unknown per node and direction,
    !! at index 3(node-1)+direction for both the equation rows and the unknown columns.
    subroutine scatter_r(ncol, enodes, hh, gg)
      integer :: ncol, enodes(9), kk, sn
      real(kind=real64) :: hh(9,3,3), gg(9,3,3)
      do il=1,3
        row=3*(ncol-1)+il
        do ik=1,3
          do kk=1,9
            sn=enodes(kk)
            col=3*(sn-1)+ik
            select case (ctype(ik,sn))
              case (0)
                A(row,col)=A(row,col)-gg(kk,il,ik)
                b(row)=b(row)-hh(kk,il,ik)*cval(ik,sn)
              case (1)
                A(row,col)=A(row,col)+hh(kk,il,ik)
                b(row)=b(row)+gg(kk,il,ik)*cval(ik,sn)
              case (10)
                A(row,col)=A(row,col)+hh(kk,il,ik)
                b(row)=b(row)+gg(kk,il,ik)*cval(ik,sn)*nfn(ik,sn)
            end select
          end do
        end do
      end do
    end subroutine scatter_r

    subroutine scatter_c(ncol, enodes, hh, gg)
      integer :: ncol, enodes(9), kk, sn
      complex(kind=real64) :: hh(9,3,3), gg(9,3,3)
      do il=1,3
        row=3*(ncol-1)+il
        do ik=1,3
          do kk=1,9
            sn=enodes(kk)
            col=3*(sn-1)+ik
            select case (ctype(ik,sn))
              case (0)
                Ac(row,col)=Ac(row,col)-gg(kk,il,ik)
                bc(row)=bc(row)-hh(kk,il,ik)*cval(ik,sn)
              case (1)
                Ac(row,col)=Ac(row,col)+hh(kk,il,ik)
                bc(row)=bc(row)+gg(kk,il,ik)*cval(ik,sn)
              case (10)
                Ac(row,col)=Ac(row,col)+hh(kk,il,ik)
                bc(row)=bc(row)+gg(kk,il,ik)*cval(ik,sn)*nfn(ik,sn)
            end select
          end do
        end do
      end do
    end subroutine scatter_c

  end subroutine do_assemble

end program ref_harness
