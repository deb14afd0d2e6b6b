*  Path-integral trajectory harness for MDynaMix (test infrastructure).
*
*  Linked with the reference's own pimd.f compiled with its PIMD entry
*  renamed to PIMDREF (preprocessor macro, source unchanged) and with the
*  reference's MPI backend (mpi_double.f): every MPI rank is one bead
*  (main.f:146-152 sets JBEAD = rank, NUMTASK = 1), and the reference main
*  program performs input, topology and start-up (main.f:139-885), then
*  calls the PIMD defined here (main.f:934), which
*    - writes this bead's start state (coordinates, internal momenta, the
*      bead-thermostat and unit constants) to "mdx_pimd_<rank>.bin",
*    - runs the reference PIMDREF for NSTEPS long steps (multiple time
*      step with NFREQ short steps, forces / NBEADS, PIFORCES springs,
*      BEADNOSE per-atom thermostat chains, PI_SCALING, GETCOM, list
*      rebuild in short step NFREQ/2+1 of every ICHNB-th long step),
*    - writes the end state, the thermostat chains and the last energies,
*      then finalises MPI.
*  Record layout as in mdx_dump.f.
*
      SUBROUTINE PIMD
      include "common_inc.f"
      real*8 DS(32)
      integer NIS(8)
      character*32 FNAM
      write(FNAM,'(a,i0,a)')'mdx_pimd_',JBEAD,'.bin'
      open(unit=78,file=FNAM,access='stream',
     +     form='unformatted',status='replace')
      NIS(1)=NSTOT
      NIS(2)=NSTEPS
      NIS(3)=ICHNB
      NIS(4)=0
      if(LPITH)NIS(4)=1
      NIS(5)=NBEADS
      NIS(6)=JNH
      NIS(7)=NFREQ
      NIS(8)=NSTTOT
      call TRJWI('PSIZES  ',8,NIS)
      DS(1)=TSTEP
      DS(2)=UNITT
      DS(3)=UNITL
      DS(4)=UNITM
      DS(5)=DT
      DS(6)=TRTEMP
      DS(7)=BTAU
      DS(8)=DQQT
      DS(9)=FNH
      DS(10)=CONVEQ
      DS(11)=BEADFAC
      DS(12)=UNITE
      DS(13)=TSTEB
      DS(14)=BOXL
      DS(15)=BOYL
      DS(16)=BOZL
      call TRJWD('PSCAL   ',16,DS)
      call TRJWD('MASSDI  ',NSTOT,MASSDI)
      call TRJST
      call PIMDREF
      call TRJST
      do J=1,JNH
        call TRJWD('THERMO  ',NSTOT,THERMO(1,J))
      end do
      DS(1)=PE1
      DS(2)=PE2
      DS(3)=PELS1
      DS(4)=PELS2
      DS(5)=SPE
      DS(6)=PINT
      DS(7)=QEK
      DS(8)=TEMP
      call TRJWD('PEND    ',8,DS)
      close(78)
      call FINAL
      END
*
*  Coordinates and velocities (reference units: Angstrom, internal
*  momentum units).
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
