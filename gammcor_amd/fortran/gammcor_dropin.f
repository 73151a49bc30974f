C     Drop-in replacements of the reference's AC-CAS routines, with the
C     reference's own names and argument lists: TwoNO1 (misc.f:166),
C     AB_CAS (interpa.f:3659), ERPASYMM1 (interpa.f:1652), ACEneERPA
C     (acfd.f:4969), ACEInteg (acfd.f:4832) and ACEInteg_FOFO
C     (abfofo.f90:3664).  Linked into the reference
C     program ahead of (or instead of) its own definitions, they read the
C     same COMMON blocks (commons.inc of the reference tree) and the same
C     rdm2.dat, and hand the work to libgammcor_hip through the C ABI.
C     The arithmetic mode comes from GAMMCOR_HIP_PREC (default FP64).
C
      Subroutine TwoNO1(TNO,URe,NBasis,NInte2)
      use gammcor_hip
      Implicit Real*8 (A-H,O-Z)
      Dimension TNO(NInte2),URe(NBasis,NBasis)
      Logical, Save :: Said = .False.
      If(.Not.Said) Write(6,'(1X,A)') 'gammcor_hip: TwoNO1 on the GPU'
      Said=.True.
      Call gmc_check(gmc_twono1(TNO,URe,NBasis,Int(NInte2,8)),'TwoNO1')
      Return
      End
C
C     The packed active 2-RDM of rdm2.dat as AB_CAS reads it
C     (interpa.f:3734-3741): a line "I J K L X" holds 2 Gamma(J,L,I,K).
      Subroutine GmcRdm2Act(RDM2Act,NAct)
      Implicit Real*8 (A-H,O-Z)
      Dimension RDM2Act(*)
      External NAddrRDM
      NRDM2Act=NAct**2*(NAct**2+1)/2
      RDM2Act(1:Max(1,NRDM2Act))=0.D0
      If(NAct.Eq.0) Return
      Open(NewUnit=IU,File='rdm2.dat',Status='Old')
   10 Read(IU,*,End=20) I,J,K,L,X
      RDM2Act(NAddrRDM(J,L,I,K,NAct))=0.5D0*X
      GoTo 10
   20 Close(IU)
      Return
      End
C
C     ACCAS runs with IndX the identity and NDim = NDimX (accas.f:53,
C     accas.f:192-195); the GPU routines assume both.
      Subroutine GmcCheckIndX(IndX,NDimX,NDim,Where)
      Implicit Real*8 (A-H,O-Z)
      Character(*) Where
      Dimension IndX(*)
      Do I=1,NDimX
      If(IndX(I).Ne.I) Then
      Write(6,'(3A)') ' gammcor_hip: ',Where,' needs IndX(i) = i'
      Stop 1
      EndIf
      EndDo
      If(NDim.Ne.NDimX) Then
      Write(6,'(3A)') ' gammcor_hip: ',Where,' needs NDim = NDimX'
      Stop 1
      EndIf
      Return
      End
C
      Subroutine AB_CAS(ABPLUS,ABMIN,ETot,URe,Occ,XOne,TwoNO,IPair,
     $ IndN,IndX,NDimX,NBasis,NDim,NInte1,NInte2,ACAlpha)
      use gammcor_hip
      Implicit Real*8 (A-H,O-Z)
      Include 'commons.inc'
      Dimension ABPLUS(NDim,NDim),ABMIN(NDim,NDim),
     $ URe(NBasis,NBasis),Occ(NBasis),XOne(NInte1),TwoNO(NInte2),
     $ IPair(NBasis,NBasis),IndN(2,NDim),IndX(NDim)
      Real*8, Allocatable :: RDM2Act(:)
      Call GmcCheckIndX(IndX,NDimX,NDim,'AB_CAS')
      Allocate(RDM2Act(Max(1,NAcCAS**2*(NAcCAS**2+1)/2)))
      Call GmcRdm2Act(RDM2Act,NAcCAS)
      Call gmc_check(gmc_ab_cas(ABPLUS,ABMIN,ETot,URe,Occ,XOne,TwoNO,
     $ Int(NInte2,8),IndN,NDimX,NBasis,NInAcCAS,NAcCAS,IGem,RDM2Act,
     $ ACAlpha,CICoef),'AB_CAS')
      Deallocate(RDM2Act)
      Return
      End
C
      Subroutine ERPASYMM1(EigVecR,Eig,APLSQRT,ABMIN,NBasis,NDimX)
      use gammcor_hip
      Implicit Real*8 (A-H,O-Z)
      Dimension APLSQRT(NDimX,NDimX),ABMIN(NDimX,NDimX),
     $ EigVecR(NDimX*NDimX),Eig(NDimX)
      NNeg=0
      Call gmc_check(gmc_erpasymm1(EigVecR,Eig,APLSQRT,ABMIN,NDimX,
     $ NNeg),'ERPASYMM1')
      If(NNeg.Ne.0) Write(6,'(1X,A,I0)')
     $ 'gammcor_hip: negative eigenvalues of A+: ',NNeg
      Return
      End
C
      Subroutine ACEneERPA(ECorr,EigVecR,Eig,TwoNO,URe,Occ,XOne,
     $ IndN,NBasis,NInte1,NInte2,NDimX,NGem)
      use gammcor_hip
      Implicit Real*8 (A-H,O-Z)
      Include 'commons.inc'
      Dimension EigVecR(NDimX*NDimX),Eig(NDimX),TwoNO(NInte2),
     $ URe(NBasis,NBasis),Occ(NBasis),XOne(NInte1),IndN(2,NDimX)
      Call gmc_check(gmc_aceneerpa(ECorr,EigVecR,Eig,TwoNO,
     $ Int(NInte2,8),Occ,IGem,CICoef,IndN,NBasis,NDimX),'ACEneERPA')
      Return
      End
C
C     The CASSCF branch of ACEInteg with NoSt = 1 (acfd.f:4884-4889,
C     4954-4965); the other branches stop with a message.
      Subroutine ACEInteg(ECorr,TwoNO,URe,Occ,XOne,UNOAO,
     $ ABPLUS,ABMIN,EigVecR,Eig,
     $ EGOne,NGOcc,
     $ Title,NBasis,NInte1,NInte2,NDim,NGem,IndAux,ACAlpha,
     $ IndN,IndX,NDimX)
      use gammcor_hip
      Implicit Real*8 (A-H,O-Z)
      Character*60 Title
      Include 'commons.inc'
      Dimension URe(NBasis,NBasis),Occ(NBasis),TwoNO(NInte2),
     $ XOne(NInte1),UNOAO(NBasis,NBasis),IndAux(NBasis),
     $ ABPLUS(NDim*NDim),ABMIN(NDim*NDim),EigVecR(NDim*NDim),
     $ Eig(NDim),EGOne(NGem),IndX(NDim),IndN(2,NDim)
      Real*8, Allocatable :: RDM2Act(:)
      Logical, Save :: Said = .False.
      If(ICASSCF.Ne.1.Or.IFlFrag1.Ne.0.Or.NoSt.Ne.1.Or.
     $ (IFunSR.Ne.0.And.IFunSRKer.Eq.1)) Then
      Write(6,'(1X,A)') 'gammcor_hip: ACEInteg supports ICASSCF=1, '//
     $ 'IFlFrag1=0, NoSt=1 and no SR kernel'
      Stop 1
      EndIf
      Call GmcCheckIndX(IndX,NDimX,NDim,'ACEInteg')
      IPrec=gmc_prec_from_env()
      If(.Not.Said) Write(6,'(1X,A,I0)')
     $ 'gammcor_hip: ACEInteg on the GPU, precision mode ',IPrec
      Said=.True.
      Allocate(RDM2Act(Max(1,NAcCAS**2*(NAcCAS**2+1)/2)))
      Call GmcRdm2Act(RDM2Act,NAcCAS)
      Call gmc_check(gmc_aceinteg(IPrec,ECorr,EGOne(1),TwoNO,
     $ Int(NInte2,8),URe,Occ,XOne,IGem,RDM2Act,IndN,NDimX,NBasis,
     $ NInAcCAS,NAcCAS,ACAlpha,ABPLUS,ABMIN,EigVecR,Eig),'ACEInteg')
      Deallocate(RDM2Act)
C     CICoef as AB_CAS leaves it in /APSG/: sqrt(n), negative for n < 1/2
      Do I=1,NBasis
      CICoef(I)=Sign(Sqrt(Occ(I)),Occ(I)-0.5D0)
      EndDo
      Return
      End
C
C     AC0CAS (acfd.f:1902-2614), the linearised AC correction, on the GPU.
C     Inputs as the reference takes them (COMMON /CAS/ and /APSG/, rdm2.dat);
C     ECorr and ETot come back, the scratch arrays ABPLUS, ABMIN, EigY and
C     Eig are not written.  fock.dat is not read: the reference only checks
C     the virtual-inactive blocks against it.
      Subroutine AC0CAS(ECorr,ETot,TwoNO,Occ,URe,XOne,
     $ ABPLUS,ABMIN,EigY,Eig,
     $ IndN,IndX,NDimX,NBasis,NDim,NInte1,NInte2)
      use gammcor_hip
      Implicit Real*8 (A-H,O-Z)
      Include 'commons.inc'
      Dimension URe(NBasis,NBasis),XOne(NInte1),Occ(NBasis),
     $ TwoNO(NInte2),ABPLUS(NDimX*NDimX),ABMIN(NDimX*NDimX),
     $ Eig(NDimX),EigY(NDimX*NDimX),IndX(NDim),IndN(2,NDim)
      Real*8, Allocatable :: RDM2Act(:)
      If(NoSt.Ne.1) Then
      Write(6,'(1X,A)') 'gammcor_hip: AC0CAS supports NoSt=1'
      Stop 1
      EndIf
      Call GmcCheckIndX(IndX,NDimX,NDim,'AC0CAS')
      Write(6,'(1X,A)') 'gammcor_hip: AC0CAS on the GPU'
      Allocate(RDM2Act(Max(1,NAcCAS**2*(NAcCAS**2+1)/2)))
      Call GmcRdm2Act(RDM2Act,NAcCAS)
      Call gmc_check(gmc_ac0cas(ECorr,ETot,TwoNO,Int(NInte2,8),URe,Occ,
     $ XOne,IGem,RDM2Act,IndN,NDimX,NBasis,NInAcCAS,NAcCAS),'AC0CAS')
      Deallocate(RDM2Act)
C     CICoef as AC0CAS leaves it in /APSG/ (acfd.f:2015-2020)
      Do I=1,NBasis
      CICoef(I)=Sign(Sqrt(Occ(I)),Occ(I)-0.5D0)
      EndDo
      Return
      End
C
C     A whole file of fixed-length records (the reference's direct-access
C     FFOO and FOFO files carry no record markers) into X(1:N).
      Subroutine GmcReadRecords(FName,X,N)
      Implicit Real*8 (A-H,O-Z)
      Character(*) FName
      Integer(8) N,ISize
      Dimension X(N)
      Inquire(File=FName,Size=ISize)
      If(ISize.Ne.8*N) Then
      Write(6,'(3A,I0,A,I0)') ' gammcor_hip: ',FName,' has ',ISize,
     $ ' bytes, expected ',8*N
      Stop 1
      EndIf
      Open(NewUnit=IU,File=FName,Access='STREAM',Form='UNFORMATTED',
     $ Status='OLD')
      Read(IU) X
      Close(IU)
      Return
      End
C
C     The CASSCF branch of ACEInteg_FOFO with NoSt = 1 (abfofo.f90:
C     3712-3719, 3767-3786).  The two integral classes are read from the
C     FFOO and FOFO record files that tran4_gen wrote (tran.f90:313-423):
C     record k+(l-1)*NOcc of FFOO holds (kl|ij) for all i,j, record
C     k+(l-1)*NBasis of FOFO holds (kl|ij) for all i and occupied j.  The
C     reference defines this routine inside MODULE abfofo, so a caller is
C     bound to this external one at link time (the module's symbol renamed
C     in the caller's object, or the module's copy removed).
      Subroutine ACEInteg_FOFO(ECorr,URe,Occ,XOne,UNOAO,
     $ ABPLUS,ABMIN,EigVecR,Eig,EGOne,NGOcc,CICoef,
     $ NBasis,NInte1,NDim,NGem,IndAux,ACAlpha,
     $ IGemIN,NAct,INActive,NELE,IndN,IndX,NDimX,
     $ NoSt,ICASSCF,IFlFrag1,IFunSR,IFunSRKer)
      use gammcor_hip
      Implicit Real*8 (A-H,O-Z)
      Dimension URe(NBasis,NBasis),Occ(NBasis),XOne(NInte1),
     $ UNOAO(NBasis,NBasis),ABPLUS(NDim*NDim),ABMIN(NDim*NDim),
     $ EigVecR(NDimX*NDimX),Eig(NDimX),EGOne(NGem),CICoef(NBasis),
     $ IndAux(NBasis),IGemIN(NBasis),IndN(2,NDim),IndX(NDim)
      Real*8, Allocatable :: RDM2Act(:)
      Real*8, Allocatable, Save :: FFOO(:),FOFO(:)
      Integer(8) NJ,NK
      Logical, Save :: Said = .False.
      If(ICASSCF.Ne.1.Or.IFlFrag1.Ne.0.Or.NoSt.Ne.1.Or.
     $ (IFunSR.Ne.0.And.IFunSRKer.Eq.1)) Then
      Write(6,'(1X,A)') 'gammcor_hip: ACEInteg_FOFO supports '//
     $ 'ICASSCF=1, IFlFrag1=0, NoSt=1 and no SR kernel'
      Stop 1
      EndIf
      Call GmcCheckIndX(IndX,NDimX,NDim,'ACEInteg_FOFO')
      NOcc=INActive+NAct
      NJ=Int(NOcc,8)**2*Int(NBasis,8)**2
      NK=NJ
      If(Allocated(FFOO)) Then
      If(Size(FFOO,Kind=8).Ne.NJ) Deallocate(FFOO,FOFO)
      EndIf
      If(.Not.Allocated(FFOO)) Allocate(FFOO(NJ),FOFO(NK))
      Call GmcReadRecords('FFOO',FFOO,NJ)
      Call GmcReadRecords('FOFO',FOFO,NK)
      IPrec=gmc_prec_from_env()
      If(.Not.Said) Write(6,'(1X,A,I0)')
     $ 'gammcor_hip: ACEInteg_FOFO on the GPU, precision mode ',IPrec
      Said=.True.
      Allocate(RDM2Act(Max(1,NAct**2*(NAct**2+1)/2)))
      Call GmcRdm2Act(RDM2Act,NAct)
      Call gmc_check(gmc_aceinteg_fofo(IPrec,ECorr,EGOne(1),FFOO,FOFO,
     $ URe,Occ,XOne,IGemIN,RDM2Act,IndN,NDimX,NBasis,INActive,NAct,
     $ ACAlpha,ABPLUS,ABMIN,EigVecR,Eig),'ACEInteg_FOFO')
      Deallocate(RDM2Act)
      Return
      End
C
C     AC0CAS_FOFO (abfofo.f90:1485-1750), the FOFO form of the linearised
C     AC correction, on the GPU.  The two integral classes come from the
C     record files IntJFile ((FF|OO)) and IntKFile ((FO|FO)) in the layout
C     ACEInteg_FOFO reads, the active 2-RDM from rdm2.dat as ABPM0_FOFO
C     reads it (abfofo.f90:1161-1168).  ECorr and ETot come back; the
C     scratch arrays ABPLUS and ABMIN are not written.  Like ACEInteg_FOFO
C     the reference defines it inside MODULE abfofo, so callers are bound to
C     this external one at link time.
      Subroutine AC0CAS_FOFO(ECorr,ETot,Occ,URe,XOne,ABPLUS,ABMIN,
     $ IndN,IndX,IGemIN,NAct,INActive,NDimX,NBasis,NDim,NInte1,
     $ NoSt,IntJFile,IntKFile)
      use gammcor_hip
      Implicit Real*8 (A-H,O-Z)
      Character(*) IntJFile,IntKFile
      Dimension Occ(NBasis),URe(NBasis,NBasis),XOne(NInte1),
     $ ABPLUS(NDimX*NDimX),ABMIN(NDimX*NDimX),IndN(2,NDim),IndX(NDim),
     $ IGemIN(NBasis)
      Real*8, Allocatable :: RDM2Act(:),FFOO(:),FOFO(:)
      Integer(8) NJ
      If(NoSt.Ne.1) Then
      Write(6,'(1X,A)') 'gammcor_hip: AC0CAS_FOFO supports NoSt=1'
      Stop 1
      EndIf
      Call GmcCheckIndX(IndX,NDimX,NDim,'AC0CAS_FOFO')
      NOcc=INActive+NAct
      NJ=Int(NOcc,8)**2*Int(NBasis,8)**2
      Allocate(FFOO(NJ),FOFO(NJ))
      Call GmcReadRecords(IntJFile,FFOO,NJ)
      Call GmcReadRecords(IntKFile,FOFO,NJ)
      Write(6,'(1X,A)') 'gammcor_hip: AC0CAS_FOFO on the GPU'
      Allocate(RDM2Act(Max(1,NAct**2*(NAct**2+1)/2)))
      Call GmcRdm2Act(RDM2Act,NAct)
      Call gmc_check(gmc_ac0cas_fofo(ECorr,ETot,FFOO,FOFO,URe,Occ,XOne,
     $ IGemIN,RDM2Act,IndN,NDimX,NBasis,INActive,NAct),'AC0CAS_FOFO')
      Deallocate(RDM2Act,FFOO,FOFO)
      Return
      End
