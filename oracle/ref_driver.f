C     Driver for the compiled reference (oracle/_ref).  Test
C     infrastructure only: it fills the reference's COMMON blocks from a
C     seeded synthetic system, runs the reference's own TwoNO1
C     (misc.f:166) and then either
C       grid  : ACEInteg (acfd.f:4832) at each alpha node read from
C               inputs.bin, printing W(alpha), ECASSCF and ERPA
C               eigenvalue summaries at full precision;
C       accas : the reference's ACCAS (accas.f:2) end to end, with its own
C               pair selection and 5-point GauLeg (acfd.f:120);
C       dumpab: AB_CAS (interpa.f:3659) at the first alpha node, writing
C               ABPLUS/ABMIN to abplus.bin/abmin.bin;
C       fofo  : the FOFO variant: the (FF|OO) and (FO|FO) classes of the
C               transformed integrals written as the FFOO and FOFO record
C               files tran4_gen produces (tran.f90:313-423), then
C               ACEInteg_FOFO (abfofo.f90:3664) at each alpha node, printing
C               as grid does;
C       ac0   : AC0CAS (acfd.f:1902), the linearised AC correction, printing
C               its ECorr and the CASSCF energy;
C       ac0fofo: its FOFO version AC0CAS_FOFO (abfofo.f90:1485) on the FFOO
C               and FOFO record files written as fofo does;
C       e2disp: SAPT's second-order dispersion energy of two monomers in
C               one basis (sapt_main.f90:2593-2714, sapt_pol.f90:1684-2034):
C               per monomer (inputs_A.bin / inputs_B.bin with rdm2_A.dat /
C               rdm2_B.dat, the AO integrals of twoel_ao.bin) its TwoNO1,
C               AB_CAS at alpha = 1, ERPASYMMXY and the response vectors
C               (C_p-C_q)(Y-X) written as PROP_A / PROP_B; the
C               intermolecular integrals come as the TWOMOAB records; then
C               the reference's e2disp, printing SAPT%e2disp;
C       formats: the reference's own readers of its input files on files
C               written by gammcor_amd.formats (no transform, no solve):
C               readtwoint with the Molpro source (sorter.f90:31) writes
C               AOTWOSORT; read_2rdm_molpro, read_1rdm_molpro,
C               read_mo_molpro and readoneint_molpro (types.f90) fill
C               arrays written to ref_rdm2.bin, ref_rdm1.bin, ref_cmo.bin
C               and ref_oneint.bin.
C     Pair selection for grid/dumpab mirrors accas.f:30-99 with
C     IFlCore=1, ThrSelAct=1e-8, ThrQVirt=1e-7 (types.f90 defaults).
      Program RefDrv
      use types
      use abfofo, only : ACEInteg_FOFO, AC0CAS_FOFO
      Implicit Real*8 (A-H,O-Z)
      Include 'commons.inc'
      Character*60 Title
      Character*16 Mode
      Real*8, Allocatable :: Occ(:),XOne(:),UAOMO(:,:),URe(:,:),
     $ TwoNO(:),AG(:),WG(:),ABPLUS(:),ABMIN(:),EigVecR(:),Eig(:),
     $ EGOne(:),UNOAO(:,:)
      Integer, Allocatable :: IndN(:,:),IndX(:),IndAux(:)
      Integer Hdr(5)
      Integer(8) IC0,IC1,ICR
      Integer(8) NAddr3
      External NAddr3
      type(SystemBlock) :: System
C
      Call Get_Command_Argument(1,Mode)
      If(Mode(1:6).Eq.'e2disp') Then
      Call RefE2Disp
      Stop
      EndIf
      Title='synthetic'
      Open(11,File='inputs.bin',Access='STREAM',Form='UNFORMATTED',
     $ Status='OLD')
      Read(11) Hdr
      NBasis=Hdr(1)
      NInAc=Hdr(2)
      NAct=Hdr(3)
      NEL=Hdr(4)
      NAl=Hdr(5)
      NInte1=NBasis*(NBasis+1)/2
      NInte2=NInte1*(NInte1+1)/2
      If(Mode(1:7).Eq.'formats') Then
      Close(11)
      Call RefFormats(NBasis,NAct)
      Stop
      EndIf
      Allocate(Occ(NBasis),XOne(NInte1),UAOMO(NBasis,NBasis),
     $ URe(NBasis,NBasis),TwoNO(NInte2),AG(NAl),WG(NAl),
     $ UNOAO(NBasis,NBasis))
      Read(11) Occ
      Read(11) XOne
      Read(11) UAOMO
      Read(11) AG
      Read(11) WG
      Close(11)
      Open(12,File='twoel_ao.bin',Access='STREAM',Form='UNFORMATTED',
     $ Status='OLD')
      Read(12) TwoNO
      Close(12)
C
C     COMMON BLOCKS
C
      XELE=Float(NEL)
      NELE=NEL
      IFunSR=0
      IFunSRKer=0
      ITwoEl=1
      ICASSCF=1
      NAcCAS=NAct
      NInAcCAS=NInAc
      NStates=1
      IFlAC=1
      IFlSnd=0
      IFlAC0D=0
      IFlAC0DP=0
      IFlCore=1
      IFlFrag1=0
      IFl12=0
      NoSt=1
      IFlACFREQ=0
      ThrSelAct=1.D-8
      ThrAct=1.D-8
      ThrQVirt=1.D-7
      NGem=3
      Do I=1,NBasis
      IGem(I)=3
      If(I.Le.NInAc+NAct) IGem(I)=2
      If(I.Le.NInAc) IGem(I)=1
      EndDo
      If(NInAc.Eq.0) Then
      NGem=2
      Do I=1,NBasis
      IGem(I)=2
      If(I.Le.NAct) IGem(I)=1
      EndDo
      EndIf
C
      Call System_Clock(IC0,ICR)
      Call TwoNO1(TwoNO,UAOMO,NBasis,NInte2)
      Call System_Clock(IC1)
      Write(6,'(A,ES15.6)') 'TIME_TWONO1 ',Dble(IC1-IC0)/Dble(ICR)
      S1=0.D0
      S2=0.D0
      Do I=1,NInte2
      S1=S1+TwoNO(I)
      S2=S2+TwoNO(I)**2
      EndDo
      Write(6,'(A,2ES25.17)') 'TWONO_SUMS ',S1,S2
      Do I=1,NBasis,7
      J=Max(1,I/2)
      K=Min(NBasis,I+3)
      L=Max(1,K-5)
      Write(6,'(A,4I6,ES25.17)') 'TWONO ',I,J,K,L,
     $ TwoNO(NAddr3(I,J,K,L))
      EndDo
C
      Do I=1,NBasis
      Do J=1,NBasis
      URe(I,J)=0.D0
      UNOAO(I,J)=UAOMO(I,J)
      EndDo
      URe(I,I)=1.D0
      EndDo
C
      If(Mode(1:5).Eq.'accas') Then
      ENuc=0.D0
      Call ACCAS(ETot,ENuc,TwoNO,URe,UNOAO,Occ,XOne,
     $ Title,NBasis,NInte1,NInte2,NGem,System)
      Stop
      EndIf
C
C     PAIR SELECTION (accas.f:30-99)
C
      NDim=NBasis*(NBasis-1)/2
      Allocate(IndN(2,NDim),IndX(NDim),IndAux(NBasis))
      Do I=1,NBasis
      IndAux(I)=2
      If(I.Le.NELE) IndAux(I)=0
      If(Occ(I).Lt.1.D0.And.Occ(I).Ne.0.D0) IndAux(I)=1
      EndDo
      Ind=0
      Do I=1,NBasis
      Do J=1,I-1
      IS=IndAux(I)+IndAux(J)
      If(IS.Ne.0.And.IS.Ne.4) Then
      If(IndAux(I).Eq.1.And.IndAux(J).Eq.1.And.
     $ Abs(Occ(I)-Occ(J))/Occ(I).Lt.ThrSelAct) Then
      Continue
      Else
      If(Abs(Occ(I)+Occ(J)-2.D0).Gt.1.D-10.And.
     $ Abs(Occ(I)+Occ(J)).Gt.ThrQVirt) Then
      Ind=Ind+1
      IndX(Ind)=Ind
      IndN(1,Ind)=I
      IndN(2,Ind)=J
      EndIf
      EndIf
      EndIf
      EndDo
      EndDo
      NDimX=Ind
      Write(6,'(A,I8)') 'NDIMX ',NDimX
C
      Allocate(ABPLUS(NDimX*NDimX),ABMIN(NDimX*NDimX),
     $ EigVecR(NDimX*NDimX),Eig(NDimX),EGOne(NGem))
C
      If(Trim(Mode).Eq.'ac0') Then
C     AC0CAS reads fock.dat for a check of its virtual-inactive blocks
C     unless IDALTON is set (acfd.f:2365-2375, 2399-2402); the check
C     only prints
      IDALTON=1
      Call AC0CAS(ECorr,ETot,TwoNO,Occ,URe,XOne,
     $ ABPLUS,ABMIN,EigVecR,Eig,
     $ IndN,IndX,NDimX,NBasis,NDimX,NInte1,NInte2)
      Write(6,'(A,ES25.17)') 'AC0ECORR ',ECorr
      Write(6,'(A,ES25.17)') 'ECASSCF ',ETot
      Stop
      EndIf
C
      If(Mode(1:6).Eq.'dumpab') Then
      Call AB_CAS(ABPLUS,ABMIN,ECASSCF,URe,Occ,XOne,TwoNO,IndAux,
     $ IndN,IndX,NDimX,NBasis,NDimX,NInte1,NInte2,AG(1))
      Write(6,'(A,ES25.17)') 'ECASSCF ',ECASSCF
      Open(13,File='abplus.bin',Access='STREAM',Form='UNFORMATTED')
      Write(13) ABPLUS
      Close(13)
      Open(13,File='abmin.bin',Access='STREAM',Form='UNFORMATTED')
      Write(13) ABMIN
      Close(13)
      Stop
      EndIf
C
      If(Mode(1:4).Eq.'fofo'.Or.Mode(1:7).Eq.'ac0fofo') Then
      NOcc=NInAc+NAct
      Open(14,File='FFOO',Status='REPLACE',Access='DIRECT',
     $ Form='UNFORMATTED',Recl=8*NBasis*NBasis)
      Do L=1,NOcc
      Do K=1,NOcc
      Write(14,Rec=K+(L-1)*NOcc) ((TwoNO(NAddr3(K,L,I,J)),I=1,NBasis),
     $ J=1,NBasis)
      EndDo
      EndDo
      Close(14)
      Open(14,File='FOFO',Status='REPLACE',Access='DIRECT',
     $ Form='UNFORMATTED',Recl=8*NBasis*NOcc)
      Do L=1,NOcc
      Do K=1,NBasis
      Write(14,Rec=K+(L-1)*NBasis) ((TwoNO(NAddr3(K,L,I,J)),I=1,NBasis),
     $ J=1,NOcc)
      EndDo
      EndDo
      Close(14)
      EndIf
C
      If(Mode(1:7).Eq.'ac0fofo') Then
      Call AC0CAS_FOFO(ECorr,ETot,Occ,URe,XOne,ABPLUS,ABMIN,
     $ IndN,IndX,IGem,NAct,NInAc,NDimX,NBasis,NDimX,NInte1,
     $ NoSt,'FFOO','FOFO')
      Write(6,'(A,ES25.17)') 'AC0ECORR ',ECorr
      Write(6,'(A,ES25.17)') 'ECASSCF ',ETot
      Stop
      EndIf
C
      ECorr=0.D0
      Do IA=1,NAl
      Call System_Clock(IC0,ICR)
      If(Mode(1:4).Eq.'fofo') Then
      Call ACEInteg_FOFO(ECorrA,URe,Occ,XOne,UNOAO,
     $ ABPLUS,ABMIN,EigVecR,Eig,EGOne,NGOcc,CICoef,
     $ NBasis,NInte1,NDimX,NGem,IndAux,AG(IA),
     $ IGem,NAcCAS,NInAcCAS,NELE,IndN,IndX,NDimX,
     $ NoSt,ICASSCF,IFlFrag1,IFunSR,IFunSRKer)
      Else
      Call ACEInteg(ECorrA,TwoNO,URe,Occ,XOne,UNOAO,
     $ ABPLUS,ABMIN,EigVecR,Eig,
     $ EGOne,NGOcc,
     $ Title,NBasis,NInte1,NInte2,NDimX,NGem,IndAux,AG(IA),
     $ IndN,IndX,NDimX)
      EndIf
      Call System_Clock(IC1)
      Write(6,'(A,I4,ES15.6)') 'TIME_ACEINTEG ',IA,
     $ Dble(IC1-IC0)/Dble(ICR)
      ECorr=ECorr+WG(IA)*ECorrA
      ES=0.D0
      NPos=0
      Do K=1,NDimX
      If(Eig(K).Gt.1.D-3.And.Eig(K).Lt.1.D8) Then
      ES=ES+Eig(K)
      NPos=NPos+1
      EndIf
      EndDo
      Write(6,'(A,I4,3ES25.17)') 'WALPHA ',IA,AG(IA),WG(IA),ECorrA
      Write(6,'(A,I4,I8,4ES25.17)') 'EIG ',IA,NPos,ES,Eig(1),
     $ Eig(NDimX/2),Eig(NDimX)
      Write(6,'(A,ES25.17)') 'ECASSCF ',EGOne(1)
      EndDo
      Write(6,'(A,ES25.17)') 'ECORR ',ECorr
C
      End

      Subroutine RefFormats(NBasis,NAct)
C     The reference's readers on the files of gammcor_amd.formats; their
C     results written as raw doubles for the test to compare.
      use types
      use sorter
      Implicit None
      Integer NBasis,NAct,IWarn,NTr
      Real*8, Allocatable :: TwoRDM(:),OneRDM(:),CMO(:,:),XMone(:)
      NTr=NBasis*(NBasis+1)/2
      Allocate(TwoRDM(NAct**2*(NAct**2+1)/2),OneRDM(NTr),
     $ CMO(NBasis,NBasis),XMone(NTr))
      Call readtwoint(NBasis,2,'AOTWOINT.mol','AOTWOSORT')
      IWarn=0
      Call read_2rdm_molpro(TwoRDM,1,0,'RDM.mol',IWarn,NAct)
      Call read_1rdm_molpro(OneRDM,1,0,'RDM.mol',IWarn,NBasis)
      Call read_mo_molpro(CMO,'MOLPRO.MOPUN','CASORB  ',NBasis)
      Call readoneint_molpro(XMone,'AOONEINT.mol','ONEHAMIL',.True.,
     $ NTr)
      Open(20,File='ref_rdm2.bin',Access='STREAM',Form='UNFORMATTED',
     $ Status='REPLACE')
      Write(20) TwoRDM
      Close(20)
      Open(20,File='ref_rdm1.bin',Access='STREAM',Form='UNFORMATTED',
     $ Status='REPLACE')
      Write(20) OneRDM
      Close(20)
      Open(20,File='ref_cmo.bin',Access='STREAM',Form='UNFORMATTED',
     $ Status='REPLACE')
      Write(20) CMO
      Close(20)
      Open(20,File='ref_oneint.bin',Access='STREAM',Form='UNFORMATTED',
     $ Status='REPLACE')
      Write(20) XMone
      Close(20)
      Write(6,'(A,I4)') 'FORMATS_OK ',IWarn
      End

      Subroutine RefE2Disp
C     Two monomers through the reference's SAPT response chain and its
C     E2disp (sapt_pol.f90:1684).  The branch of e2disp for
C     ICASSCF = 0, ISERPA = 0 forms the same coupled sum by plain loops
C     and needs no alpha=0 response files (its uncoupled part is skipped).
      use types
      use sapt_pol
      Implicit Real*8 (A-H,O-Z)
      Include 'commons.inc'
      type(SystemBlock) :: Mon(2)
      type(FlagsData) :: Flags
      type(SaptData) :: SAPT
      Character*1 Tag(2)
      Character*80 Line
      Integer Hdr(5)
      Real*8, Allocatable :: Occ(:),XOne(:),UAOMO(:,:),URe(:,:),
     $ AO(:),TwoNO(:),ABPLUS(:),ABMIN(:),EigY(:),EigX(:),Eig(:),
     $ EVec(:),AG(:),WG(:)
      Integer, Allocatable :: IndN(:,:),IndX(:),IndAux(:)
      Data Tag /'A','B'/
      Do IM=1,2
      Open(11,File='inputs_'//Tag(IM)//'.bin',Access='STREAM',
     $ Form='UNFORMATTED',Status='OLD')
      Read(11) Hdr
      NBasis=Hdr(1)
      NInAc=Hdr(2)
      NAct=Hdr(3)
      NEL=Hdr(4)
      NAl=Hdr(5)
      NInte1=NBasis*(NBasis+1)/2
      NInte2=NInte1*(NInte1+1)/2
      Allocate(Occ(NBasis),XOne(NInte1),UAOMO(NBasis,NBasis),
     $ URe(NBasis,NBasis),AG(NAl),WG(NAl),TwoNO(NInte2))
      Read(11) Occ
      Read(11) XOne
      Read(11) UAOMO
      Read(11) AG
      Read(11) WG
      Close(11)
      Open(12,File='twoel_ao.bin',Access='STREAM',Form='UNFORMATTED',
     $ Status='OLD')
      Read(12) TwoNO
      Close(12)
C     this monomer's rdm2.dat, where AB_CAS reads it
      Open(13,File='rdm2_'//Tag(IM)//'.dat',Status='OLD')
      Open(14,File='rdm2.dat',Status='REPLACE')
   10 Read(13,'(A)',End=20) Line
      Write(14,'(A)') Trim(Line)
      GoTo 10
   20 Close(13)
      Close(14)
      XELE=Float(NEL)
      NELE=NEL
      IFunSR=0
      IFunSRKer=0
      ITwoEl=1
      ICASSCF=1
      NAcCAS=NAct
      NInAcCAS=NInAc
      NoSt=1
      ThrSelAct=1.D-8
      ThrQVirt=1.D-7
      Do I=1,NBasis
      IGem(I)=3
      If(I.Le.NInAc+NAct) IGem(I)=2
      If(I.Le.NInAc) IGem(I)=1
      EndDo
      If(NInAc.Eq.0) Then
      Do I=1,NBasis
      IGem(I)=2
      If(I.Le.NAct) IGem(I)=1
      EndDo
      EndIf
      Call TwoNO1(TwoNO,UAOMO,NBasis,NInte2)
      URe=0.D0
      Do I=1,NBasis
      URe(I,I)=1.D0
      EndDo
C     pair selection as the grid mode (accas.f:30-99)
      NDim=NBasis*(NBasis-1)/2
      Allocate(IndN(2,NDim),IndX(NDim),IndAux(NBasis))
      Do I=1,NBasis
      IndAux(I)=2
      If(I.Le.NELE) IndAux(I)=0
      If(Occ(I).Lt.1.D0.And.Occ(I).Ne.0.D0) IndAux(I)=1
      EndDo
      Ind=0
      Do I=1,NBasis
      Do J=1,I-1
      IS=IndAux(I)+IndAux(J)
      If(IS.Ne.0.And.IS.Ne.4) Then
      If(.Not.(IndAux(I).Eq.1.And.IndAux(J).Eq.1.And.
     $ Abs(Occ(I)-Occ(J))/Occ(I).Lt.ThrSelAct)) Then
      If(Abs(Occ(I)+Occ(J)-2.D0).Gt.1.D-10.And.
     $ Abs(Occ(I)+Occ(J)).Gt.ThrQVirt) Then
      Ind=Ind+1
      IndX(Ind)=Ind
      IndN(1,Ind)=I
      IndN(2,Ind)=J
      EndIf
      EndIf
      EndIf
      EndDo
      EndDo
      NDimX=Ind
      Allocate(ABPLUS(NDimX*NDimX),ABMIN(NDimX*NDimX),
     $ EigY(NDimX*NDimX),EigX(NDimX*NDimX),Eig(NDimX),
     $ EVec(NDimX*NDimX))
C     sapt_main.f90:2604-2660: A+-B at alpha = 1, ERPASYMMXY, and
C     EigVecR = (C_p - C_q)(Y - X)
      Call AB_CAS(ABPLUS,ABMIN,ETot,URe,Occ,XOne,TwoNO,IndAux,
     $ IndN,IndX,NDimX,NBasis,NDimX,NInte1,NInte2,1.D0)
      Call ERPASYMMXY(EigY,EigX,Eig,ABPLUS,ABMIN,Occ,IndN,NDimX,
     $ NBasis)
      Do J=1,NDimX
      Do I=1,NDimX
      IP=IndN(1,I)
      IQ=IndN(2,I)
      EVec((J-1)*NDimX+I)=(CICoef(IP)-CICoef(IQ))
     $ *(EigY((J-1)*NDimX+I)-EigX((J-1)*NDimX+I))
      EndDo
      EndDo
C     the response file as writeresp writes it (sapt_main.f90:4068)
      Open(15,File='PROP_'//Tag(IM),Form='UNFORMATTED',
     $ Status='REPLACE')
      Write(15) EVec
      Write(15) Eig
      Close(15)
      Mon(IM)%NBasis=NBasis
      Mon(IM)%NDimX=NDimX
      Mon(IM)%num0=NInAc
      Mon(IM)%num1=NAct
      Mon(IM)%num2=NBasis-NInAc-NAct
      Allocate(Mon(IM)%IndN(2,NDimX),Mon(IM)%CICoef(NBasis))
      Mon(IM)%IndN(1:2,1:NDimX)=IndN(1:2,1:NDimX)
      Mon(IM)%CICoef(1:NBasis)=CICoef(1:NBasis)
      Write(6,'(A,A,I8,ES25.17)') 'MONOMER ',Tag(IM),NDimX,ETot
      Deallocate(Occ,XOne,UAOMO,URe,AG,WG,TwoNO,IndN,IndX,IndAux,
     $ ABPLUS,ABMIN,EigY,EigX,Eig,EVec)
      EndDo
      Flags%ICASSCF=0
      Flags%ISERPA=0
      SAPT%SemiCoupled=.False.
      SAPT%Wexcit=.False.
      SAPT%IPrint=0
      Call e2disp(Flags,Mon(1),Mon(2),SAPT)
      Write(6,'(A,ES25.17)') 'E2DISP ',SAPT%e2disp
      End
