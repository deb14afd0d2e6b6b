*  Trajectory harness for the MDynaMix integrators (test infrastructure).
*
*  Linked with the reference's own mdstep.f compiled with its SULTAN and
*  DOUBLE entries renamed to SULTREF and DOUBREF (preprocessor macros,
*  source unchanged), so the reference main program performs input,
*  topology and start-up (main.f:139-885) and calls the SULTAN or DOUBLE
*  defined here (main.f:918-926: SULTAN with constraints, DOUBLE without),
*  which
*    - writes the start state (coordinates, velocities, masses, step and
*      thermostat constants) to "mdx_traj.bin",
*    - for SULTAN sets NSTTOT = 1 so the integrator's first-step velocity
*      rescaling (mdstep.f:126-134, MSTEP = 1 only) does not apply,
*    - runs the reference integrator for NSTEPS (long) steps: SULTREF
*      (leap-frog, SHEJK constraints, SCALING(3) thermostat, GETCOM, list
*      rebuild every ICHNB steps, ALLFORCE) or DOUBREF (multiple time
*      step with NFREQ short steps, SCALING(1)/(2), GETCOM, SLFORCES and
*      FFORCES, list rebuild at short step NFREQ/2+1 of every ICHNB-th
*      long step),
*    - writes the end state and the last step's energies.
*  Record layout as in mdx_dump.f.
*
      SUBROUTINE SULTAN
      call TRJRUN(0)
      END
*
      SUBROUTINE DOUBLE
      call TRJRUN(1)
      END
*
      SUBROUTINE TRJRUN(MODE)
      include "common_inc.f"
      integer MODE
      real*8 DS(32)
      integer NIS(8)
      open(unit=78,file='mdx_traj.bin',access='stream',
     +     form='unformatted',status='replace')
      if(MODE.eq.0)NSTTOT=1
      NIS(1)=NSTOT
      NIS(2)=NSTEPS
      NIS(3)=ICHNB
      NIS(4)=0
      if(LNVT)NIS(4)=1
      NIS(5)=0
      if(LSHEJK)NIS(5)=1
      NIS(6)=MODE
      NIS(7)=NFREQ
      NIS(8)=NSTTOT
      call TRJWI('TSIZES  ',8,NIS)
      DS(1)=TSTEP
      DS(2)=UNITT
      DS(3)=UNITL
      DS(4)=UNITM
      DS(5)=DT
      DS(6)=TRTEMP
      DS(7)=QT
      DS(8)=DQT
      DS(9)=FNST
      DS(10)=SC
      DS(11)=TOLER
      DS(12)=BDTOL
      DS(13)=BOXL
      DS(14)=BOYL
      DS(15)=BOZL
      DS(16)=UNITE
      DS(17)=TSTEB
      call TRJWD('TSCAL   ',17,DS)
      call TRJWD('MASSDI  ',NSTOT,MASSDI)
      call TRJST
      if(MODE.eq.0)then
        call SULTREF
      else
        call DOUBREF
      end if
      call TRJST
      DS(1)=SC
      DS(2)=TEMP
      DS(3)=TKE
      DS(4)=PE1
      DS(5)=PE2
      DS(6)=PELS1
      DS(7)=PELS2
      DS(8)=SPE
      DS(9)=PINT
      call TRJWD('TEND    ',9,DS)
      close(78)
      stop
      END
*
*  Coordinates and velocities (reference units: Angstrom, internal
*  momentum units of SULTAN).
      SUBROUTINE TRJST
      include "common_inc.f"
      call TRJWD('SX      ',NSTOT,SX)
      call TRJWD('SY      ',NSTOT,SY)
      call TRJWD('SZ      ',NSTOT,SZ)
      call TRJWD('VX      ',NSTOT,VX)
      call TRJWD('VY      ',NSTOT,VY)
      call TRJWD('VZ      ',NSTOT,VZ)
      END
*
      SUBROUTINE TRJWI(NAM,N,IA)
      character*8 NAM
      integer N,IA(*)
      write(78)NAM,1,N
      if(N.gt.0)write(78)(IA(I),I=1,N)
      END
      SUBROUTINE TRJWD(NAM,N,A)
      character*8 NAM
      integer N
      real*8 A(*)
      write(78)NAM,2,N
      if(N.gt.0)write(78)(A(I),I=1,N)
      END
