C     TEST INFRASTRUCTURE ONLY (the oracle).  Never linked into the product.
C
C     Driver that replaces the main program of the reference solver
C     (multall-open-18.3.f:89-142) so that the reference's own
C     NEW_READIN, SETUP and LOOP can be run without a terminal and their
C     COMMON state captured:
C       - unit 1 (/dev/tty in the reference, mf:91) is a plain file that
C         answers 'Y' to the negative-volume prompt (mf:7310-7318);
C       - after SETUP the raw bytes of every COMMON block are written to
C         'setup.bin';
C       - LOOP then runs to NSTEPS_MAX and ends with STOP (mf:5451); an
C         atexit hook registered in ref_dump.c writes 'final.bin'.
C     The block base addresses and sizes come from LOC() of the first and
C     last member of each block as declared in commall-open-18.3.
C
      PROGRAM REFHARNESS
      INCLUDE 'commall-open-18.3'
      CHARACTER*1 ANSIN
C
      OPEN(UNIT=1,FILE='tty_answers')
      WRITE(1,*) 'Y'
      REWIND(1)
      OPEN(UNIT=4,FILE ='stage.log')
      OPEN(UNIT=7,FILE ='flow_out',   FORM = 'unformatted')
      OPEN(UNIT=11,FILE='global.plt', FORM = 'unformatted')
      OPEN(UNIT=3,FILE ='results.out')
      OPEN(UNIT=12,FILE='stopit')
      WRITE(12,*) 0
      CLOSE(12)
C
      ANSIN = 'N'
      CALL NEW_READIN
      CALL SETUP(ANSIN)
C
      CALL MHBEGIN(KIND(1.0))
      CALL MHBLK('BKUTIL',LOC(NAME),LOC(RADDEG)+KIND(1.0))
      CALL MHBLK('BKINTG',LOC(IM),LOC(IN_PRESS)+4)
      CALL MHBLK('BKGEOM',LOC(X),LOC(ANGL_DWN2(KD))+KIND(1.0))
      CALL MHBLK('BKAREA',LOC(ABX),LOC(ABTOT(ID,JD,KD))+KIND(1.0))
      CALL MHBLK('BKFLUX',LOC(XFLUX),LOC(SOURCE(ID,JD,KD))+KIND(1.0))
      CALL MHBLK('BKFLOW',LOC(FLOWX),LOC(SPECFLOW)+KIND(1.0))
      CALL MHBLK('BKSEC',LOC(VX),LOC(WT(ID,JD,KD))+KIND(1.0))
      CALL MHBLK('BKPRIM',LOC(RO),LOC(ROWT(ID,JD,KD))+KIND(1.0))
      CALL MHBLK('BKFFLO',LOC(IN_FLOW),LOC(FLOW(JD))+KIND(1.0))
      CALL MHBLK('BKFRAC',LOC(FR),LOC(QSPAN(KD))+KIND(1.0))
      CALL MHBLK('BKSMTH',LOC(SFTIN),LOC(F3)+KIND(1.0))
      CALL MHBLK('BKLOSS',LOC(FRACPB),
     &           LOC(DPDS_CELL(ID,JD,KD))+KIND(1.0))
      CALL MHBLK('BKINPT',LOC(PO1),LOC(F_PDOWN)+KIND(1.0))
      CALL MHBLK('BKMGRD',LOC(IR),LOC(AVG_BLK(NRS))+KIND(1.0))
      CALL MHBLK('BKGAS',LOC(CP),LOC(GA1GAS)+KIND(1.0))
      CALL MHBLK('BKCNTL',LOC(CFL),LOC(ROELIMIT(NRS))+KIND(1.0))
      CALL MHBLK('BKSTAG',LOC(NROWS),LOC(UBLADE(JD,KD))+KIND(1.0))
      CALL MHBLK('BKTIP',LOC(KTIPS),LOC(FRACTIP2(NRS))+KIND(1.0))
      CALL MHBLK('BKDIFF',LOC(DRO),LOC(DRORVT(ID,JD,KD))+KIND(1.0))
      CALL MHBLK('BKDIST',LOC(FTL),LOC(FBR(ID,KD))+KIND(1.0))
      CALL MHBLK('BKNEWLOS',LOC(DIST_MIN),
     &           LOC(YPLUS_IM(JD,KD))+KIND(1.0))
      CALL MHBLK('BKSPALOS',LOC(TRANS_KVIS),
     &           LOC(SPAREVAR(ID,JD,KD))+KIND(1.0))
      CALL MHBLK('BKOUT',LOC(PO_OUT_AVG),LOC(PITCH_IN_MID)+KIND(1.0))
      CALL MHBLK('BKTFLOW',LOC(SFPBLD),LOC(S_DIST(JD))+KIND(1.0))
      CALL MHBLK('BKATOT',LOC(ATOTK1),LOC(ATOTIM(JD,KD))+KIND(1.0))
      CALL MHDUMP('setup.bin')
C
      CALL MHATEXIT('final.bin')
      CALL LOOP
      STOP
      END
