*  mdx_forces.f -- drop-in replacement of the reference's forces.f.
*
*  Linked in place of forces.o, this file provides the entry points the
*  reference driver calls (CHCNB, ALLFORCE, FFORCES, SLFORCES from
*  mdstep.f/pimd.f; ELRCLJ from main.f:816) with their names, arguments and
*  COMMON side effects, and delegates the force path to the MI355X engine
*  through the C ABI of include/mdx.h (ISO_C_BINDING).
*
*  On the first CHCNB the topology and parameters held in COMMON
*  (/ADDREFS/, /SITTES/, /LISTNB/ INBB, /BONBON/, /ANGANG/, /TORTOR/,
*  /SIZES/, /ENERGY/ COULF, /MDEE/) are copied to 0-based arrays and the
*  engine is created with reference list semantics (MDX_LIST_REFERENCE).
*  Every call passes SX/SY/SZ and the molecular centres X/Y/Z and writes
*  FX/GX/HX and the energy/virial scalars the integrators and averages
*  read (ZEROFS/CM_ADD* semantics, service.f:397-453, scalar.f:28-116).
*  A nonzero engine status is reported and ends the run through FINAL,
*  as the reference does on its own errors (forces.f:2188-2194).
*  MDX_PRECISION=fp64 in the environment selects the engine's FP64 mode
*  (MDX_PRECISION_FP64) instead of the default mixed-precision path.
*
*  With Average_int yes (ISTAV > 0, input.f:372-374) the fast-group calls
*  also fill the per-template bonded sums BR/BE, AR/AE, TR/TE that GETAVR
*  averages (aver.f:43-112), from mdx_get_terms.
*
*  Not supported here (the engine returns MDX_ERR_ARG or the shim stops):
*  reaction field, hexagonal/octahedral cells, fixed molecules, restraints
*  and external fields of PULLBACK (forces.f:1756-1799).
*
      module mdxc
      use iso_c_binding
      implicit none
      integer, parameter :: MAXT = 16, MAXTP = MAXT*(MAXT+1)/2+1
      type, bind(C) :: mdx_topology_t
        integer(c_int32_t) :: natoms, nmols, nmoltypes, nljtypes
        integer(c_int32_t) :: ntypepairs
        type(c_ptr) :: atom_moltype, atom_mol, atom_ljtype
        type(c_ptr) :: atom_charge, atom_charge0, atom_mass
        type(c_ptr) :: type_ee, type_move, type_c14el, type_c14lj
        type(c_ptr) :: type_pair
        type(c_ptr) :: a6, b12, a6_0, b12_0, a6_14, b12_14
        type(c_ptr) :: excl_start, excl_atom, excl_is14
        integer(c_int32_t) :: nbonds
        type(c_ptr) :: bond_i, bond_j, bond_param
        integer(c_int32_t) :: nbondparams
        type(c_ptr) :: bp_morse, bp_r0, bp_k, bp_d, bp_rho
        integer(c_int32_t) :: nangles
        type(c_ptr) :: angle_i, angle_j, angle_k, angle_param
        integer(c_int32_t) :: nangleparams
        type(c_ptr) :: ap_theta0, ap_k
        integer(c_int32_t) :: ntors
        type(c_ptr) :: tors_i, tors_j, tors_k, tors_l, tors_param
        integer(c_int32_t) :: ntorsparams
        type(c_ptr) :: tp_kind, tp_mult, tp_k, tp_delta
        type(c_ptr) :: tp_k1, tp_k2, tp_k3, tp_k4, tp_k5
        integer(c_int32_t) :: nspc
        type(c_ptr) :: spc_atom, spc_par, spc_bond
      end type
      type, bind(C) :: mdx_params_t
        real(c_double) :: box(3), rcut, rshort, skin, alpha, fexp
        real(c_double) :: coulomb_factor
        integer(c_int32_t) :: list_mode, ee_enabled
        real(c_double) :: ee_raex, ee_self_scale
        integer(c_int32_t) :: precision
      end type
      type, bind(C) :: mdx_energies_t
        real(c_double) :: pe1, pe2, pels1, pels2, spe, pint
        real(c_double) :: vir1, vir2, virb, virx, viry, virz
        real(c_double) :: virfx, virfy, virfz, wirs, wirss
        integer(c_int64_t) :: npairs_short, npairs_long
        real(c_double) :: pote1(MAXTP), potl1(MAXTP)
        real(c_double) :: pes141(MAXT), psr141(MAXT), spee(MAXT)
        real(c_double) :: qek, virpi
      end type
      interface
        integer(c_int) function mdx_create(t, p, h) bind(C)
          import :: c_int, c_ptr, mdx_topology_t, mdx_params_t
          type(mdx_topology_t), intent(in) :: t
          type(mdx_params_t), intent(in) :: p
          type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function mdx_set_positions(h, x, y, z) bind(C)
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), intent(in) :: x(*), y(*), z(*)
        end function
        integer(c_int) function mdx_set_box(h, b) bind(C)
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), intent(in) :: b(3)
        end function
        integer(c_int) function mdx_set_molecule_com(h, x, y, z)
     +    bind(C)
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double), intent(in) :: x(*), y(*), z(*)
        end function
        integer(c_int) function mdx_build_list(h) bind(C)
          import :: c_int, c_ptr
          type(c_ptr), value :: h
        end function
        integer(c_int) function mdx_list_counts(h, ns, nl) bind(C)
          import :: c_int, c_ptr, c_int64_t
          type(c_ptr), value :: h
          integer(c_int64_t), intent(out) :: ns, nl
        end function
        integer(c_int) function mdx_get_list(h, s1, s2, cs, l1, l2, cl)
     +    bind(C)
          import :: c_int, c_ptr, c_int32_t, c_int64_t
          type(c_ptr), value :: h
          integer(c_int32_t) :: s1(*), s2(*), l1(*), l2(*)
          integer(c_int64_t), value :: cs, cl
        end function
        integer(c_int) function mdx_compute(h, grp, want, fx, fy, fz,
     +    gx, gy, gz, hx, hy, hz, e) bind(C)
          import :: c_int, c_ptr, c_int32_t, c_double, mdx_energies_t
          type(c_ptr), value :: h
          integer(c_int32_t), value :: grp, want
          real(c_double) :: fx(*), fy(*), fz(*), gx(*), gy(*), gz(*)
          real(c_double) :: hx(*), hy(*), hz(*)
          type(mdx_energies_t), intent(out) :: e
        end function
*  the same entry point with only F given (G and H not wanted)
        integer(c_int) function mdx_compute_f(h, grp, want, fx, fy, fz,
     +    gx, gy, gz, hx, hy, hz, e) bind(C, name='mdx_compute')
          import :: c_int, c_ptr, c_int32_t, c_double, mdx_energies_t
          type(c_ptr), value :: h
          integer(c_int32_t), value :: grp, want
          real(c_double) :: fx(*), fy(*), fz(*)
          type(c_ptr), value :: gx, gy, gz, hx, hy, hz
          type(mdx_energies_t), intent(out) :: e
        end function
        integer(c_int) function mdx_get_terms(h, br, be, ar, ae, tr,
     +    te) bind(C)
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: h
          real(c_double) :: br(*), be(*), ar(*), ae(*), tr(*), te(*)
        end function
        type(c_ptr) function mdx_last_error(h) bind(C)
          import :: c_ptr
          type(c_ptr), value :: h
        end function
      end interface
*  engine handle and the 0-based copies the topology struct points at
      type(c_ptr), save :: hdl = c_null_ptr
      logical, save :: created = .false.
      real(c_double), save :: lastbox(3) = 0.d0
      type(mdx_topology_t), save, target :: topo
      type(mdx_params_t), save, target :: prm
      integer(c_int32_t), allocatable, target, save :: amtyp(:),
     +  amol(:), aljt(:), tee(:), tmov(:), tpair(:), exs(:), exa(:),
     +  bi(:), bj(:), bp(:), bmo(:), ai(:), aj(:), ak(:), ap(:),
     +  ti(:), tj(:), tk(:), tl(:), tp(:), tkind(:), tmult(:),
     +  sat(:), sbd(:)
      integer(c_int8_t), allocatable, target, save :: ex14(:)
      real(c_double), allocatable, target, save :: aq(:), aq0(:),
     +  am(:), c14e(:), c14l(:), ta6(:), tb12(:), ta60(:), tb120(:),
     +  ta614(:), tb1214(:), br0(:), bk(:), bd(:), brho(:), ath(:),
     +  akk(:), tk0(:), tdel(:), tk1(:), tk2(:), tk3(:), tk4(:),
     +  tk5(:), spar(:)
      end module mdxc
*
*  Report an engine error and stop the run through FINAL.
      subroutine MDXERR(IST, WHERE)
      use mdxc
      integer(c_int) IST
      character*(*) WHERE
      character(len=256), pointer :: msg
      character(kind=c_char), pointer :: cm(:)
      integer K
      write(*,*)'!!! MI355X force engine error ',IST,' in ',WHERE
      if(c_associated(hdl))then
        call c_f_pointer(mdx_last_error(hdl), cm, [256])
        do K=1,256
          if(cm(K).eq.c_null_char)exit
        end do
        write(*,*)'    ',cm(1:K-1)
      end if
      call FINAL
      end
*
*  Copy COMMON topology/parameters to the engine's 0-based layout and
*  create the engine (first CHCNB).
      subroutine MDXINI
      use mdxc
      include "common_inc.f"
      integer(c_int) IST
      integer NEX, K, L, IA, JA, NTT
      character*16 PRECV
      N=NSTOT
      allocate(amtyp(N),amol(N),aljt(N),aq(N),aq0(N),am(N))
      do I=1,N
        amtyp(I)=ITYPE(I)-1
        amol(I)=NNUM(I)-1
        aljt(I)=INBT(NSITE(I))-1
        aq(I)=Q(I)
        aq0(I)=CHARGE(NSITE(I))
        am(I)=MASS(NSITE(I))
      end do
      allocate(tee(NTYPES),tmov(NTYPES),c14e(NTYPES),c14l(NTYPES))
      allocate(tpair(NTYPES*NTYPES))
      do I=1,NTYPES
        tee(I)=IEE(I)
        tmov(I)=0
        if(LMOVE(I))tmov(I)=1
        c14e(I)=C14EL(I)
        c14l(I)=C14LJ(I)
        do J=1,NTYPES
          tpair((I-1)*NTYPES+J)=MDX(I,J)-1
        end do
      end do
      NTT=NNBT*NNBT
      allocate(ta6(NTT),tb12(NTT),ta60(NTT),tb120(NTT),ta614(NTT),
     +  tb1214(NTT))
      do I=1,NNBT
        do J=1,NNBT
          K=(I-1)*NNBT+J
          ta6(K)=A6LJ(I,J)
          tb12(K)=B12LJ(I,J)
          ta60(K)=A6LJ0(I,J)
          tb120(K)=B12LJ0(I,J)
          ta614(K)=A6_14(I,J)
          tb1214(K)=B12_14(I,J)
        end do
      end do
      NEX=0
      do I=1,N
        NEX=NEX+NNBB(I)
      end do
      allocate(exs(N+1),exa(max(NEX,1)),ex14(max(NEX,1)))
      exs(1)=0
      K=0
      do I=1,N
        do L=1,NNBB(I)
          K=K+1
          exa(K)=iabs(INBB(I,L))-1
          ex14(K)=0
          if(INBB(I,L).lt.0)ex14(K)=1
        end do
        exs(I+1)=K
      end do
      allocate(bi(max(NRBT,1)),bj(max(NRBT,1)),bp(max(NRBT,1)))
      do K=1,NRBT
        bi(K)=IBI(K)-1
        bj(K)=JBJ(K)-1
        bp(K)=IBUK(K)-1
      end do
      allocate(bmo(max(NRBB,1)),br0(max(NRBB,1)),bk(max(NRBB,1)),
     +  bd(max(NRBB,1)),brho(max(NRBB,1)))
      do K=1,NRBB
        bmo(K)=0
        if(ID(K).eq.1)bmo(K)=1
        br0(K)=RB(K)
        bk(K)=FB(K)
        bd(K)=DB(K)
        brho(K)=RO(K)
      end do
      allocate(ai(max(NRAT,1)),aj(max(NRAT,1)),ak(max(NRAT,1)),
     +  ap(max(NRAT,1)),ath(max(NRAA,1)),akk(max(NRAA,1)))
      do K=1,NRAT
        ai(K)=IAI(K)-1
        aj(K)=JAJ(K)-1
        ak(K)=KAK(K)-1
        ap(K)=IAUK(K)-1
      end do
      do K=1,NRAA
        ath(K)=RA(K)
        akk(K)=FA(K)
      end do
      allocate(ti(max(NRRT,1)),tj(max(NRRT,1)),tk(max(NRRT,1)),
     +  tl(max(NRRT,1)),tp(max(NRRT,1)))
      do K=1,NRRT
        ti(K)=ITI(K)-1
        tj(K)=JTJ(K)-1
        tk(K)=KTK(K)-1
        tl(K)=LTL(K)-1
        tp(K)=ITUK(K)-1
      end do
      allocate(tkind(max(NRTT,1)),tmult(max(NRTT,1)),tk0(max(NRTT,1)),
     +  tdel(max(NRTT,1)),tk1(max(NRTT,1)),tk2(max(NRTT,1)),
     +  tk3(max(NRTT,1)),tk4(max(NRTT,1)),tk5(max(NRTT,1)))
      do K=1,NRTT
        tkind(K)=ITORS(K)
        tmult(K)=NMUL(K)
        tk0(K)=FT(K)
        tdel(K)=RT(K)
        tk1(K)=FT1(K)
        tk2(K)=FT2(K)
        tk3(K)=FT3(K)
        tk4(K)=FT4(K)
        tk5(K)=FT5(K)
      end do
*  fSPC water (IPOT 1 or 2): STRBND records, forces.f:555-605
      NSPCM=0
      do ITYP=1,NTYPES
        if(IPOT(ITYP).eq.1.or.IPOT(ITYP).eq.2)NSPCM=NSPCM+NSPEC(ITYP)
      end do
      allocate(sat(max(3*NSPCM,1)),spar(max(12*NSPCM,1)))
      allocate(sbd(max(NSPCM,1)))
      K=0
      do ITYP=1,NTYPES
        if(IPOT(ITYP).eq.1.or.IPOT(ITYP).eq.2)then
          I1=IADB(ITYP)
          do M=1,NSPEC(ITYP)
            K=K+1
            IML=ISADDR(ITYP)+(M-1)*NSITS(ITYP)
            sat(3*K-2)=IML
            sat(3*K-1)=IML+1
            sat(3*K)=IML+2
            sbd(K)=I1-1
            spar(12*K-11)=RB(I1)
            spar(12*K-10)=RB(I1+1)
            spar(12*K-9)=RB(I1+2)
            if(IPOT(ITYP).eq.2)then
              spar(12*K-8)=DB(I1)
              spar(12*K-7)=DB(I1+1)
            else
              spar(12*K-8)=FB(I1)
              spar(12*K-7)=FB(I1+1)
            end if
            spar(12*K-6)=FB(I1+2)
            spar(12*K-5)=RO(I1)
            spar(12*K-4)=RO(I1+1)
            spar(12*K-3)=-884.63D0*1.D3/AVSNO/UNITE
            spar(12*K-2)=467.31D0*1.D3/AVSNO/UNITE
            spar(12*K-1)=dble(IPOT(ITYP))
            spar(12*K)=0.d0
            if(LMOVE(ITYP))spar(12*K)=1.d0
          end do
        end if
      end do
      topo%natoms=N
      topo%nmols=NOP
      topo%nmoltypes=NTYPES
      topo%nljtypes=NNBT
      topo%ntypepairs=MOLINT
      topo%atom_moltype=c_loc(amtyp)
      topo%atom_mol=c_loc(amol)
      topo%atom_ljtype=c_loc(aljt)
      topo%atom_charge=c_loc(aq)
      topo%atom_charge0=c_loc(aq0)
      topo%atom_mass=c_loc(am)
      topo%type_ee=c_loc(tee)
      topo%type_move=c_loc(tmov)
      topo%type_c14el=c_loc(c14e)
      topo%type_c14lj=c_loc(c14l)
      topo%type_pair=c_loc(tpair)
      topo%a6=c_loc(ta6)
      topo%b12=c_loc(tb12)
      topo%a6_0=c_loc(ta60)
      topo%b12_0=c_loc(tb120)
      topo%a6_14=c_loc(ta614)
      topo%b12_14=c_loc(tb1214)
      topo%excl_start=c_loc(exs)
      topo%excl_atom=c_loc(exa)
      topo%excl_is14=c_loc(ex14)
      topo%nbonds=NRBT
      topo%bond_i=c_loc(bi)
      topo%bond_j=c_loc(bj)
      topo%bond_param=c_loc(bp)
      topo%nbondparams=NRBB
      topo%bp_morse=c_loc(bmo)
      topo%bp_r0=c_loc(br0)
      topo%bp_k=c_loc(bk)
      topo%bp_d=c_loc(bd)
      topo%bp_rho=c_loc(brho)
      topo%nangles=NRAT
      topo%angle_i=c_loc(ai)
      topo%angle_j=c_loc(aj)
      topo%angle_k=c_loc(ak)
      topo%angle_param=c_loc(ap)
      topo%nangleparams=NRAA
      topo%ap_theta0=c_loc(ath)
      topo%ap_k=c_loc(akk)
      topo%ntors=NRRT
      topo%tors_i=c_loc(ti)
      topo%tors_j=c_loc(tj)
      topo%tors_k=c_loc(tk)
      topo%tors_l=c_loc(tl)
      topo%tors_param=c_loc(tp)
      topo%ntorsparams=NRTT
      topo%tp_kind=c_loc(tkind)
      topo%tp_mult=c_loc(tmult)
      topo%tp_k=c_loc(tk0)
      topo%tp_delta=c_loc(tdel)
      topo%tp_k1=c_loc(tk1)
      topo%tp_k2=c_loc(tk2)
      topo%tp_k3=c_loc(tk3)
      topo%tp_k4=c_loc(tk4)
      topo%tp_k5=c_loc(tk5)
      topo%nspc=NSPCM
      topo%spc_atom=c_loc(sat)
      topo%spc_par=c_loc(spar)
      topo%spc_bond=c_loc(sbd)
      prm%box(1)=BOXL
      prm%box(2)=BOYL
      prm%box(3)=BOZL
      prm%rcut=RCUTT
      prm%rshort=SHORT
      prm%skin=0.d0
      prm%alpha=ALPHAD
      prm%fexp=FEXP
      prm%coulomb_factor=COULF
      prm%list_mode=0
      prm%ee_enabled=0
      prm%ee_raex=0.d0
      prm%ee_self_scale=1.d0
C  MDX_PRECISION=fp64 in the environment: the engine's FP64 mode
      prm%precision=0
      call get_environment_variable('MDX_PRECISION',PRECV)
      if(PRECV(1:4).eq.'fp64'.or.PRECV(1:4).eq.'FP64')prm%precision=1
      if(LMDEE)then
        prm%ee_enabled=1
        prm%ee_raex=RAEX
        prm%ee_self_scale=EC(ME)
      end if
      if(LHEX.or.LOCT.or.LMNBL)then
        write(*,*)'!!! GPU force engine: rectangular cell, Ewald only'
        call FINAL
      end if
      IST=mdx_create(topo,prm,hdl)
      if(IST.ne.0)call MDXERR(IST,'mdx_create')
      lastbox=prm%box
      created=.true.
      end
*
*  Pass the current box (NPT) and coordinates.
      subroutine MDXPOS
      use mdxc
      include "common_inc.f"
      integer(c_int) IST
      real(c_double) B(3)
      B(1)=BOXL
      B(2)=BOYL
      B(3)=BOZL
      if(B(1).ne.lastbox(1).or.B(2).ne.lastbox(2).or.
     +   B(3).ne.lastbox(3))then
        IST=mdx_set_box(hdl,B)
        if(IST.ne.0)call MDXERR(IST,'mdx_set_box')
        lastbox=B
        IST=mdx_build_list(hdl)
        if(IST.ne.0)call MDXERR(IST,'mdx_build_list')
      end if
      IST=mdx_set_positions(hdl,SX,SY,SZ)
      if(IST.ne.0)call MDXERR(IST,'mdx_set_positions')
      IST=mdx_set_molecule_com(hdl,X,Y,Z)
      if(IST.ne.0)call MDXERR(IST,'mdx_set_molecule_com')
      end
*
*  CHCNB (forces.f:1813): neighbour lists.  MBSH/MBLN receive the pair
*  counts; the lists themselves are exported into NBS1/NBS2/NBL1/NBL2
*  only for expanded-ensemble runs, whose DIFEN reads them (mdee.f:397).
      subroutine CHCNB(LRDF)
      use mdxc
      include "common_inc.f"
      logical LRDF
      integer(c_int) IST
      integer(c_int64_t) NPS8, NPL8
      timev0=cputime(0.d0)
      if(.not.created)call MDXINI
      IST=mdx_set_positions(hdl,SX,SY,SZ)
      if(IST.ne.0)call MDXERR(IST,'mdx_set_positions')
      IST=mdx_build_list(hdl)
      if(IST.ne.0)call MDXERR(IST,'mdx_build_list')
      IST=mdx_list_counts(hdl,NPS8,NPL8)
      if(IST.ne.0)call MDXERR(IST,'mdx_list_counts')
      MBSH=NPS8
      MBLN=NPL8
      if(LMDEE)then
        if(NPS8.gt.NBSMAX.or.NPL8.gt.NBLMAX)then
          write(*,*)' Insrease NBLMAX/NBSMAX for this job'
          call FINAL
        end if
        IST=mdx_get_list(hdl,NBS1,NBS2,int(NBSMAX,c_int64_t),
     +    NBL1,NBL2,int(NBLMAX,c_int64_t))
        if(IST.ne.0)call MDXERR(IST,'mdx_get_list')
      end if
      timev=timev+cputime(timev0)
      if(IPRINN.ge.7)write(*,*)' Task ',TASKID,
     +'total num.of interactions: ',MBSH,' and ',MBLN
      end
*
*  Energies of one engine call into COMMON (the scalars ZEROFS(1) and/or
*  ZEROFS(2) reset; service.f:405-445).
      subroutine MDXPUT(E, IGRP)
      use mdxc
      include "common_inc.f"
      type(mdx_energies_t) E
      integer IGRP
      do I=1,MOLINT
        POTE1(I)=E%pote1(I)
        POTL1(I)=E%potl1(I)
      end do
      do I=1,NTYPES
        PES141(I)=E%pes141(I)
        PSR141(I)=E%psr141(I)
      end do
      if(IGRP.ne.2)then
        PELS2=E%pels2
        PE2=E%pe2
        PINT=E%pint
        VIRB=E%virb
        VIR2=E%vir2
        WIRSS=E%wirss
        VIRFX=E%virfx
        VIRFY=E%virfy
        VIRFZ=E%virfz
      end if
      if(IGRP.ne.1)then
        PE1=E%pe1
        PELS1=E%pels1
        SPE=E%spe
        VIR1=E%vir1
        VIRX=E%virx
        VIRY=E%viry
        VIRZ=E%virz
        WIRS=E%wirs
        do I=1,NTYPES
          SPEE(I)=E%spee(I)
        end do
      end if
      end
*
*  Options of a fast-group call: the energy breakdown always, the
*  per-template bonded sums when GETAVR averages them (ISTAV > 0).
      integer(c_int32_t) function MDXWANT()
      use mdxc
      include "common_inc.f"
      MDXWANT=1
      if(ISTAV.gt.0)MDXWANT=3
      end
*
*  BR/BE, AR/AE, TR/TE as GETBND, STRBND, GETANG and GETTRS leave them
*  (forces.f:119-182, 655-663, 232-273, 336-447), when averaged.
      subroutine MDXTRM
      use mdxc
      include "common_inc.f"
      integer(c_int) IST
      if(ISTAV.le.0)return
      IST=mdx_get_terms(hdl,BR,BE,AR,AE,TR,TE)
      if(IST.ne.0)call MDXERR(IST,'mdx_get_terms')
      end
*
*  Restraints and external fields of PULLBACK are not part of this build.
      subroutine MDXNOPULL
      include "common_inc.f"
      if(LDRAG.or.(LRR.and.NPULL.gt.0).or.IEXT.eq.1)then
        write(*,*)'!!! GPU force engine: PULLBACK terms not supported'
        call FINAL
      end if
      end
*
*  ALLFORCE (forces.f:77-100).
      subroutine ALLFORCE
      use mdxc
      include "common_inc.f"
      type(mdx_energies_t) E
      integer(c_int) IST
      integer(c_int32_t) MDXWANT
      call ZEROAV
      call MDXPOS
      IST=mdx_compute(hdl,0,MDXWANT(),FX,FY,FZ,GX,GY,GZ,HX,HY,HZ,E)
      if(IST.ne.0)call MDXERR(IST,'mdx_compute(ALL)')
      call MDXPUT(E,0)
      call MDXTRM
*  CM_ADDLA accumulations (scalar.f:82-106)
      do I=1,MOLINT
        POTES(I)=POTES(I)+POTE1(I)
        POTLJ(I)=POTLJ(I)+POTL1(I)
      end do
      do I=1,NTYPES
        PES14(I)=PES14(I)+PES141(I)
        PSR14(I)=PSR14(I)+PSR141(I)
        SELFPE(I)=SELFPE(I)+SPEE(I)
      end do
      do M=1,NRBB
        BB(M)=BB(M)+BR(M)
        EB(M)=EB(M)+BE(M)
      end do
      do M=1,NRAA
        AA(M)=AA(M)+AR(M)
        EA(M)=EA(M)+AE(M)
      end do
      do M=1,NRTT
        TT(M)=TT(M)+TR(M)
        ET(M)=ET(M)+TE(M)
      end do
      call MDXNOPULL
      call SCLFRC(3)
      end
*
*  FFORCES (forces.f:36-51).
      subroutine FFORCES(LAVER)
      use mdxc
      include "common_inc.f"
      logical LAVER
      type(mdx_energies_t) E
      integer(c_int) IST
      integer(c_int32_t) MDXWANT
      call MDXPOS
      IST=mdx_compute_f(hdl,1,MDXWANT(),HX,HY,HZ,c_null_ptr,c_null_ptr,
     +  c_null_ptr,c_null_ptr,c_null_ptr,c_null_ptr,E)
      if(IST.ne.0)call MDXERR(IST,'mdx_compute(FAST)')
      call MDXPUT(E,1)
      call MDXTRM
      call CM_ADDF(LAVER)
      call SCLFRC(1)
      end
*
*  SLFORCES (forces.f:55-73).
      subroutine SLFORCES
      use mdxc
      include "common_inc.f"
      type(mdx_energies_t) E
      integer(c_int) IST
      call ZEROAV
      call MDXPOS
      IST=mdx_compute_f(hdl,2,1,GX,GY,GZ,c_null_ptr,c_null_ptr,
     +  c_null_ptr,c_null_ptr,c_null_ptr,c_null_ptr,E)
      if(IST.ne.0)call MDXERR(IST,'mdx_compute(SLOW)')
      call MDXPUT(E,2)
*  CM_ADDLF accumulations without its WIRS sum (scalar.f:59-71); the
*  engine returns WIRS already.
      do I=1,MOLINT
        POTES(I)=POTES(I)+POTE1(I)
        POTLJ(I)=POTLJ(I)+POTL1(I)
      end do
      do I=1,NTYPES
        PES14(I)=PES14(I)+PES141(I)
        PSR14(I)=PSR14(I)+PSR141(I)
        SELFPE(I)=SELFPE(I)+SPEE(I)
      end do
      call MDXNOPULL
      call SCLFRC(2)
      end
*
*  ELRCLJ (forces.f:1684-1746): Lennard-Jones tail corrections EPD/VIRD.
*  For E = (A6 + B12 r^-6) r^-6 beyond RC the tail per pair of sites is
*  4 pi / V * (A6/(3 RC^3) + B12/(9 RC^9)) (energy) and
*  4 pi / V * (2 A6/RC^3 + 4 B12/(3 RC^9)) (virial, sign as VRLRC).
      subroutine ELRCLJ
      include "common_inc.f"
      dimension PEL0(NTPP),VRL0(NTPP)
      logical INIT
      data INIT/.false./
      if(.not.INIT)then
        do I=1,MOLINT
          PEL0(I)=0.d0
          VRL0(I)=0.d0
        end do
        RC0=dsqrt(RSSQ)
        do ITYP=1,NTYPES
          do IS=ISADR(ITYP)+1,ISADR(ITYP+1)
            IST=INBT(IS)
            do JTYP=ITYP,NTYPES
              RC=RC0
              if(IEE(ITYP)*IEE(JTYP).eq.0)RC=RC0+RAEX
              MT=MDX(ITYP,JTYP)
              FNP=dfloat(NSPEC(ITYP))*dfloat(NSPEC(JTYP))
              if(ITYP.eq.JTYP)FNP=0.5d0*FNP
              do JS=ISADR(JTYP)+1,ISADR(JTYP+1)
                JST=INBT(JS)
                A6=A6LJ(IST,JST)
                B12=B12LJ(IST,JST)
                if(B12.gt.1.d-50.and.A6.lt.-1.d-50)then
                  R3=RC**3
                  PEL0(MT)=PEL0(MT)+(A6/(3.d0*R3)+B12/(9.d0*R3**3))*FNP
                  VRL0(MT)=VRL0(MT)+(2.d0*A6/R3+4.d0*B12/(3.d0*R3**3))
     +              *FNP
                end if
              end do
            end do
          end do
        end do
        INIT=.true.
      end if
      VIRD=0.d0
      EPD=0.d0
      do I=1,MOLINT
        PELRC(I)=4.d0*PI*PEL0(I)/VOL
        VRLRC(I)=4.d0*PI*VRL0(I)/VOL
        EPD=EPD+PELRC(I)
        VIRD=VIRD+VRLRC(I)
      end do
      end
