*  Golden-vector harness for the MDynaMix force path (test infrastructure).
*
*  Linked in place of the reference's mdstep.f, so the reference's own
*  main program performs input, topology, units, start-up and Ewald set-up
*  (main.f:139-885) and then calls SULTAN or DOUBLE (main.f:919-932).
*  Those two entry points are redefined here: instead of integrating they
*  build the neighbour list (CHCNB, forces.f:1813), evaluate the forces the
*  way the integrators do on their first step (mdstep.f:42-47 for SULTAN,
*  mdstep.f:299-307 for DOUBLE) and write every input and output of the
*  force path to the binary file "mdx_dump.bin".
*
*  Record layout (stream access): 8-character name, int32 kind
*  (1 = int32, 2 = real*8), int32 count, then count values.
*
*  Expanded-ensemble runs also record the reference's DIFFUR/DIFEN energy
*  differences toward every subensemble at each dumped call.
*
*  SULTAN additionally displaces the atoms by a deterministic pattern and
*  calls ALLFORCE again WITHOUT rebuilding the list, so the dump also pins
*  the stale-list semantics (pairs evaluated without a cutoff test).
*
      SUBROUTINE SULTAN
      include "common_inc.f"
      call MDXTIM
      open(unit=77,file='mdx_dump.bin',access='stream',
     +     form='unformatted',status='replace')
      call CHCNB(.false.)
      call ALLFORCE
      call MDXSTA(1)
      call MDXLST
      call MDXFRC
      do I=1,NSTOT
        SX(I)=SX(I)+0.03d0*sin(0.7d0*I)
        SY(I)=SY(I)+0.03d0*cos(1.3d0*I)
        SZ(I)=SZ(I)+0.03d0*sin(2.9d0*I+0.5d0)
      end do
      call ALLFORCE
      call MDXFRC
      close(77)
      stop
      END
*
      SUBROUTINE DOUBLE
      include "common_inc.f"
      call MDXTIM
      open(unit=77,file='mdx_dump.bin',access='stream',
     +     form='unformatted',status='replace')
      call CHCNB(.false.)
      call GETCOM
      call SLFORCES
      call FFORCES(.true.)
      call MDXSTA(2)
      call MDXLST
      call MDXFRC
      close(77)
      stop
      END
*
*  Timing mode (CPU baseline), set by MDX_TIME_STRIPES=N in the
*  environment.  N > 0 (serial build): the run takes the share of rank 0
*  of N replicated-data ranks (the reference's own striding: CHCNB outer
*  atoms, forces.f:1825-1830; FURIR k vectors, forces.f:1273-1277; ETERMS,
*  forces.f:1501-1502).  N = 0 (MPI build): the ranks of the MPI job.
*  Times CHCNB and one ALLFORCE (with its reductions) with the wall clock
*  between barriers; rank 0 prints
*  "MDXTIME ranks MBSH MBLN t_list t_force" (its own list lengths); then
*  the program ends.  No dump is written.
      SUBROUTINE MDXTIM
      include "common_inc.f"
      character*16 ENVV
      integer LENV,NSTR,IST
      integer*8 IC0,IC1,IC2,IRATE
      call GET_ENVIRONMENT_VARIABLE('MDX_TIME_STRIPES',ENVV,LENV,IST)
      if(IST.ne.0.or.LENV.le.0)return
      read(ENVV,*)NSTR
      if(NSTR.gt.0)then
        NUMTASK=NSTR
        TASKID=0
      end if
      call BARRIER
      call SYSTEM_CLOCK(IC0,IRATE)
      call CHCNB(.false.)
      call SYSTEM_CLOCK(IC1)
      call ALLFORCE
      call BARRIER
      call SYSTEM_CLOCK(IC2)
      if(TASKID.eq.0)write(*,'(a,3i12,2f14.6)')' MDXTIME ',NUMTASK,
     +  MBSH,MBLN,dble(IC1-IC0)/dble(IRATE),dble(IC2-IC1)/dble(IRATE)
      call FINAL
      END
*
*  Unused integrator pieces of mdstep.f: never reached from this harness.
      SUBROUTINE SCALING(IAL)
      stop 'SCALING not available in the dump harness'
      END
      SUBROUTINE SHEJK(IMOL)
      stop 'SHEJK not available in the dump harness'
      END
*
*  --- record writers ---------------------------------------------------
      SUBROUTINE MDXWI(NAM,N,IA)
      character*8 NAM
      integer N,IA(*)
      write(77)NAM,1,N
      if(N.gt.0)write(77)(IA(I),I=1,N)
      END
      SUBROUTINE MDXWD(NAM,N,A)
      character*8 NAM
      integer N
      real*8 A(*)
      write(77)NAM,2,N
      if(N.gt.0)write(77)(A(I),I=1,N)
      END
*
*  Static set-up: topology, parameters and unit factors (once).
      SUBROUTINE MDXSTA(IMODE)
      include "common_inc.f"
      integer IW(4*NTOT)
      real*8 DW(NTOT),DS(64)
      integer NIS(16)
      NIS(1)=IMODE
      NIS(2)=NSTOT
      NIS(3)=NTYPES
      NIS(4)=NOP
      NIS(5)=NNBT
      NIS(6)=NKV
      NIS(7)=NRBT
      NIS(8)=NRAT
      NIS(9)=NRRT
      NIS(10)=MOLINT
      NIS(11)=ICELL
      NIS(12)=0
      if(LMDEE)NIS(12)=1
      NIS(13)=IPRINT
      NIS(14)=NSITES
      call MDXWI('SIZES   ',14,NIS)
      DS(1)=BOXL
      DS(2)=BOYL
      DS(3)=BOZL
      DS(4)=RCUTT
      DS(5)=RSSQ
      DS(6)=SHORT
      DS(7)=SHORT2
      DS(8)=ALPHAD
      DS(9)=FEXP
      DS(10)=COULF
      DS(11)=ENERF
      DS(12)=UNITE
      DS(13)=EFACT
      DS(14)=VOL
      DS(15)=RAEX
      DS(16)=SPEI
      DS(17)=EPD
      DS(18)=VIRD
      DS(19)=UNITM
      DS(20)=UNITT
      DS(21)=UNITP
      DS(22)=TOTMAS
      DS(23)=DT
      call MDXWD('SCALARS ',23,DS)
      call MDXWI('ITYPE   ',NSTOT,ITYPE)
      call MDXWI('NNUM    ',NSTOT,NNUM)
      call MDXWI('NSITE   ',NSTOT,NSITE)
      do I=1,NSTOT
        IW(I)=INBT(NSITE(I))
      end do
      call MDXWI('INBTA   ',NSTOT,IW)
      do I=1,NSTOT
        DW(I)=CHARGE(NSITE(I))
      end do
      call MDXWD('CHARGEA ',NSTOT,DW)
      call MDXWD('Q       ',NSTOT,Q)
      do I=1,NSTOT
        DW(I)=MASSD(NSITE(I))
      end do
      call MDXWD('MASSDA  ',NSTOT,DW)
      do I=1,NSTOT
        DW(I)=MASS(NSITE(I))
      end do
      call MDXWD('MASSA   ',NSTOT,DW)
      do I=1,NTYPES
        IW(I)=IEE(I)
      end do
      call MDXWI('IEE     ',NTYPES,IW)
      do I=1,NTYPES
        IW(I)=0
        if(LMOVE(I))IW(I)=1
      end do
      call MDXWI('LMOVE   ',NTYPES,IW)
      call MDXWD('C14EL   ',NTYPES,C14EL)
      call MDXWD('C14LJ   ',NTYPES,C14LJ)
      call MDXWI('IPOT    ',NTYPES,IPOT)
      call MDXWI('IADB    ',NTYPES,IADB)
      call MDXWI('NSPEC   ',NTYPES,NSPEC)
      call MDXWI('NSITS   ',NTYPES,NSITS)
      call MDXWD('SPEEI   ',NTYPES,SPEEI)
      do I=1,NTYPES
        do J=1,NTYPES
          IW((I-1)*NTYPES+J)=MDX(I,J)
        end do
      end do
      call MDXWI('MDX     ',NTYPES*NTYPES,IW)
      call MDXTAB('A6LJ    ',A6LJ)
      call MDXTAB('B12LJ   ',B12LJ)
      call MDXTAB('A6LJ0   ',A6LJ0)
      call MDXTAB('B12LJ0  ',B12LJ0)
      call MDXTAB('A6_14   ',A6_14)
      call MDXTAB('B12_14  ',B12_14)
*  exclusion lists (BNDLST, setup.f:753-888)
      call MDXWI('NNBB    ',NSTOT,NNBB)
      K=0
      do I=1,NSTOT
        do J=1,NNBB(I)
          K=K+1
          IW(K)=INBB(I,J)
          if(K.ge.4*NTOT)stop 'exclusion dump overflow'
        end do
      end do
      call MDXWI('INBB    ',K,IW)
*  k-vectors (INIFUR, forces.f:1356)
      call MDXWD('RKX     ',NKV,RKX)
      call MDXWD('RKY     ',NKV,RKY)
      call MDXWD('RKZ     ',NKV,RKZ)
*  bonded term lists and parameters (IUNITS, setup.f:283-381)
      call MDXWI('IBI     ',NRBT,IBI)
      call MDXWI('JBJ     ',NRBT,JBJ)
      call MDXWI('IBUK    ',NRBT,IBUK)
      call MDXWI('ID      ',NRBB,ID)
      call MDXWD('RB      ',NRBB,RB)
      call MDXWD('FB      ',NRBB,FB)
      call MDXWD('DB      ',NRBB,DB)
      call MDXWD('RO      ',NRBB,RO)
      call MDXWI('IAI     ',NRAT,IAI)
      call MDXWI('JAJ     ',NRAT,JAJ)
      call MDXWI('KAK     ',NRAT,KAK)
      call MDXWI('IAUK    ',NRAT,IAUK)
      call MDXWD('RA      ',NRAA,RA)
      call MDXWD('FA      ',NRAA,FA)
      call MDXWI('ITI     ',NRRT,ITI)
      call MDXWI('JTJ     ',NRRT,JTJ)
      call MDXWI('KTK     ',NRRT,KTK)
      call MDXWI('LTL     ',NRRT,LTL)
      call MDXWI('ITUK    ',NRRT,ITUK)
      call MDXWI('ITORS   ',NRTT,ITORS)
      call MDXWI('NMUL    ',NRTT,NMUL)
      call MDXWD('FT      ',NRTT,FT)
      call MDXWD('RT      ',NRTT,RT)
      call MDXWD('FT1     ',NRTT,FT1)
      call MDXWD('FT2     ',NRTT,FT2)
      call MDXWD('FT3     ',NRTT,FT3)
      call MDXWD('FT4     ',NRTT,FT4)
      call MDXWD('FT5     ',NRTT,FT5)
      END
*
*  One NNBT x NNBT block of an LJ table (row-major by first index).
      SUBROUTINE MDXTAB(NAM,T)
      include "common_inc.f"
      character*8 NAM
      real*8 T(NNBTM,NNBTM)
      real*8 DW(NNBTM*NNBTM/4)
      if(NNBT*NNBT.gt.NNBTM*NNBTM/4)stop 'table dump overflow'
      do I=1,NNBT
        do J=1,NNBT
          DW((I-1)*NNBT+J)=T(I,J)
        end do
      end do
      call MDXWD(NAM,NNBT*NNBT,DW)
      END
*
*  Neighbour lists as built by CHCNB.
      SUBROUTINE MDXLST
      include "common_inc.f"
      integer NIS(2)
      NIS(1)=MBSH
      NIS(2)=MBLN
      call MDXWI('LISTLEN ',2,NIS)
      call MDXWI('NBS1    ',MBSH,NBS1)
      call MDXWI('NBS2    ',MBSH,NBS2)
      call MDXWI('NBL1    ',MBLN,NBL1)
      call MDXWI('NBL2    ',MBLN,NBL2)
      END
*
*  Positions, forces and every energy/virial scalar of one force call.
      SUBROUTINE MDXFRC
      include "common_inc.f"
      real*8 DS(64)
      call MDXWD('SX      ',NSTOT,SX)
      call MDXWD('SY      ',NSTOT,SY)
      call MDXWD('SZ      ',NSTOT,SZ)
      call MDXWD('XCOM    ',NOP,X)
      call MDXWD('YCOM    ',NOP,Y)
      call MDXWD('ZCOM    ',NOP,Z)
      call MDXWD('FX      ',NSTOT,FX)
      call MDXWD('FY      ',NSTOT,FY)
      call MDXWD('FZ      ',NSTOT,FZ)
      call MDXWD('GX      ',NSTOT,GX)
      call MDXWD('GY      ',NSTOT,GY)
      call MDXWD('GZ      ',NSTOT,GZ)
      call MDXWD('HX      ',NSTOT,HX)
      call MDXWD('HY      ',NSTOT,HY)
      call MDXWD('HZ      ',NSTOT,HZ)
      DS(1)=PE1
      DS(2)=PE2
      DS(3)=PELS1
      DS(4)=PELS2
      DS(5)=SPE
      DS(6)=PINT
      DS(7)=VIR1
      DS(8)=VIR2
      DS(9)=VIRB
      DS(10)=VIRX
      DS(11)=VIRY
      DS(12)=VIRZ
      DS(13)=VIRFX
      DS(14)=VIRFY
      DS(15)=VIRFZ
      DS(16)=WIRS
      DS(17)=WIRSS
      call MDXWD('ENERGY  ',17,DS)
      call MDXWD('POTE1   ',MOLINT,POTE1)
      call MDXWD('POTL1   ',MOLINT,POTL1)
      call MDXWD('PES141  ',NTYPES,PES141)
      call MDXWD('PSR141  ',NTYPES,PSR141)
      call MDXWD('SPEE    ',NTYPES,SPEE)
      call MDXWD('BR      ',NRBB,BR)
      call MDXWD('BE      ',NRBB,BE)
      call MDXWD('AR      ',NRAA,AR)
      call MDXWD('AE      ',NRAA,AE)
      call MDXWD('TR      ',NRTT,TR)
      call MDXWD('TE      ',NRTT,TE)
*  expanded ensemble: the reference's own energy differences toward every
*  subensemble ME1 (DIFFUR, DIFEN; mdee.f:265-534), with the current lists
*  and coordinates: EFN, EFO, EEN, EEO per ME1; then ME, PW1, PW2, BEE and
*  the subensembles' EC
      if(LMDEE)then
        do ME1=1,NE
          call DIFFUR(ME1,EFN,EFO)
          call DIFEN(ME1,EEN,EEO)
          DS(4*ME1-3)=EFN
          DS(4*ME1-2)=EFO
          DS(4*ME1-1)=EEN
          DS(4*ME1)=EEO
        end do
        call MDXWD('EEDIFF  ',4*NE,DS)
        DS(1)=ME
        DS(2)=PW1
        DS(3)=PW2
        DS(4)=BEE
        call MDXWD('EEPAR   ',4,DS)
        call MDXWD('EC      ',NE,EC)
      end if
      END
