! Fortran host of the MI355X pconf path: inputs, determinant space, outputs.
!
! Restates the host side of pconf that feeds the hot path, in Fortran like
! the reference:
!   ReadConfInp/ReadConfigurations/ReadCiIn   src/conf_init.f90:37-313
!   Init (CONF.DAT, orbital mapping)          src/pconf.F90:410-681
!   Rint (CONF.INT)                           src/integrals.f90:13-120
!   Input (CONF.GNT)                          src/pconf.F90:382-406
!   Kl = 2 / Ksig rule, Eps (CONF.DAT)        src/pconf.F90:345-356, 663-667
!   RintS (SGC.CON), FormD                    src/pconf.F90:683-762, 852-880
!   Dinit / Ndet / Det_List / sort_det        src/determinants.f90:47-437, 811-836
!   calcNd0                                   src/determinants.f90:22-45
!   Wdet (CONF.DET)                           src/determinants.f90:448-462
!   WriteMatrixCSR (pCONF.HIJ), ReadMatrixCSR src/matrix_io.f90:689-939
!   WriteFinalXIJ (CONF.XIJ)                  src/pconf.F90:2566-2602
!   PrintEnergies                             src/pconf.F90:3152-3242
! and hands the arrays to the device through pci_gpu_api.  Library routines
! return a status instead of stopping, so the module can also be driven from
! C or Python (the bind(C) entry points at the end).
Module pconf_host
    Use, Intrinsic :: iso_c_binding
    Use, Intrinsic :: iso_fortran_env, Only : dp => real64, int64, int32, real32
    Use pci_gpu_api
    Implicit None

    Integer, Parameter :: IP6 = 470, IPS = 600
    Real(dp), Parameter :: DPRy = 109737.315709_dp

    ! CONF.INP header
    Character(Len=16) :: confname = ' '
    Real(dp) :: Z = 0, Am = 0, XJ_av = 0, Jm = 0, Crt4 = 1.d-4, Gj = 0
    Integer :: Nso = 0, Nc = 0, Kv = 4, Nlv = 1, Ne = 0, Kl4 = 1, Nc4 = 0, N_it = 20, Kbrt = 0
    ! ci.in
    Integer :: Kl = 0, Kw = 0, KXIJ = 10, MaxNd0 = 3000, kCSF = 0, Ksig = 0, Kdsig = 0
    ! MBPT one-electron effective integrals (Kl = 2, Ksig = 1)
    Integer :: NmaxS = 0, LmaxS = 0, NhintS = 0
    Integer, Allocatable :: Iint1S(:)
    Real(dp), Allocatable :: Rsig(:), Dsig(:), Esig(:)
    Real(dp), Allocatable, Target :: DiagS(:)
    Real(dp) :: Eps(IPS) = 0, E_0 = 0
    Logical :: have_e0 = .false.
    ! MBPT with screening (Ksig = 2): SGC.CON's screening coefficients, SCRC.CON
    Real(dp) :: Scr(10) = 1
    Integer :: Ksym = 1, KmaxS = 0, Nsum = 0, NgintS = 0
    Integer, Allocatable :: Iint2S(:), Iint3S(:)
    Real(dp), Allocatable :: Rint2S(:), Dint2S(:), Eint2S(:)
    ! basis and configurations
    Integer :: Ns = 0, Nsu = 0, Nsp = 0, Nst = 0, Npos = 0, Mj = 0
    Integer :: Nn(IPS) = 0, Ll(IPS) = 0, Kk(IPS) = 0, Jj(IPS) = 0, Nf0(IPS) = 0
    Real(dp), Allocatable :: Qnl(:)
    Integer, Allocatable :: Nip(:), Nq(:), Nq0(:), Nvc(:), Nc0(:), Jz(:), Nh(:), Ndc(:), Iarr(:,:)
    Integer :: Nd = 0, Nd0 = 0, Nc1 = 0
    ! CSF basis (kCSF > 0): class of each configuration (nonequiv_conf)
    Integer :: Nconf_neq = 0
    Integer, Allocatable, Target :: Nc_neq(:)
    ! integrals
    Real(dp) :: Ecore = 0
    Integer :: Nhint = 0, Nx = 440, Nlx = 5, Ngaunt = 0
    Integer(int64) :: Ngint = 0
    Real(dp), Allocatable :: Rint1(:), Rint2(:,:), Gnt(:)
    Integer, Allocatable :: Iint1(:), Iint2(:), Iint3(:), Gin(:)
    Character(Len=512) :: host_err = ' '

Contains

    ! ------------------------------------------------------------------ utils
    Function upper(s) Result(u)
        Character(Len=*), Intent(In) :: s
        Character(Len=Len(s)) :: u
        Integer :: i, c
        u = s
        Do i = 1, Len(s)
            c = IChar(s(i:i))
            If (c >= IChar('a') .and. c <= IChar('z')) u(i:i) = Char(c - 32)
        End Do
    End Function upper

    Integer Function fail(msg)
        Character(Len=*), Intent(In) :: msg
        host_err = msg
        fail = 1
    End Function fail

    ! Electrons held by a CONF.INP entry (src/conf_init.f90:290-296)
    Integer Function entry_electrons(x)
        Real(dp), Intent(In) :: x
        Real(dp) :: a
        a = Abs(x) + 1.d-9
        entry_electrons = Int(10000*a) - 100*Int(100*a)
    End Function entry_electrons

    ! --------------------------------------------------------------- CONF.INP
    Integer Function read_conf_inp(path) Result(rc)
        Character(Len=*), Intent(In) :: path
        Character(Len=200) :: line, key, val
        Real(dp) :: q(6)
        Real(dp), Allocatable :: tmp(:)
        Integer :: u, ios, ieq, ih, i, n, ic, ne0, cap
        Logical :: pending

        rc = 0
        Open(newunit=u, file=path, status='OLD', action='READ', iostat=ios)
        If (ios /= 0) Then
            rc = fail('cannot open '//Trim(path))
            Return
        End If
        Read(u, '(A)') line
        confname = line(2:17)
        ! defaults (src/conf_init.f90:219-231)
        Kbrt = 0; Gj = 0; N_it = 20; Kl4 = 1; Nso = 0; Nc4 = 0; Kv = 4
        pending = .false.
        Do
            Read(u, '(A)', iostat=ios) line
            If (ios /= 0) Then
                rc = fail('CONF.INP ends inside the header')
                Close(u)
                Return
            End If
            ieq = Index(line, '=')
            If (ieq == 0) Then
                pending = Len_Trim(line) > 0
                Exit
            End If
            key = upper(AdjustL(line(1:ieq-1)))
            val = line(ieq+1:)
            ih = Index(val, '#')
            If (ih > 0) val = val(1:ih-1)
            Select Case (key(1:4))
            Case ('Z   '); Read(val, *) Z
            Case ('AM  '); Read(val, *) Am
            Case ('J   '); Read(val, *) XJ_av
            Case ('JM  '); Read(val, *) Jm
            Case ('NSO '); Read(val, *) Nso
            Case ('NC  '); Read(val, *) Nc
            Case ('KV  '); Read(val, *) Kv
            Case ('NLV '); Read(val, *) Nlv
            Case ('NE  '); Read(val, *) Ne
            Case ('KL4 '); Read(val, *) Kl4
            Case ('NC4 '); Read(val, *) Nc4
            Case ('GJ  '); Read(val, *) Gj
            Case ('CRT4'); Read(val, *) Crt4
            Case ('N_IT'); Read(val, *) N_it
            Case ('KBRT'); Read(val, *) Kbrt
            End Select
        End Do
        If (Nc <= 0 .or. Ne <= 0) Then
            rc = fail('CONF.INP: Nc and Ne must be positive')
            Close(u)
            Return
        End If
        cap = Nso + 16*Nc + 16
        Allocate(tmp(cap))
        n = 0
        ! core shells
        Do While (n < Nso)
            If (pending) Then
                pending = .false.
            Else
                Read(u, '(A)', iostat=ios) line
                If (ios /= 0) Exit
            End If
            Read(line, '(6(4X,F7.4))') q
            Do i = 1, 6
                If (n < Nso) Then
                    n = n + 1
                    tmp(n) = q(i)
                End If
            End Do
        End Do
        ! configurations: each starts on a new line and runs until Ne electrons
        Do ic = 1, Nc
            ne0 = 0
            Do While (ne0 < Ne)
                If (pending) Then
                    pending = .false.
                Else
                    Read(u, '(A)', iostat=ios) line
                    If (ios /= 0) Then
                        rc = fail('CONF.INP: end of file inside the configuration list')
                        Close(u)
                        Return
                    End If
                End If
                Read(line, '(6(4X,F7.4))') q
                Do i = 1, 6
                    If (Abs(q(i)) + 1.d-9 < 1.d-8) Exit
                    If (n + 1 > cap) Then
                        rc = fail('CONF.INP: configuration list too long')
                        Close(u)
                        Return
                    End If
                    n = n + 1
                    tmp(n) = q(i)
                    ne0 = ne0 + entry_electrons(q(i))
                End Do
            End Do
            If (ne0 > Ne) Then
                rc = fail('CONF.INP: too many electrons in a configuration')
                Close(u)
                Return
            End If
        End Do
        Close(u)
        Nsp = n
        If (Allocated(Qnl)) Deallocate(Qnl)
        Allocate(Qnl(Nsp))
        Qnl = tmp(1:Nsp)
        If (Nc4 <= 0 .or. Nc4 > Nc) Nc4 = Nc
    End Function read_conf_inp

    Subroutine read_ci_in(path)
        Character(Len=*), Intent(In) :: path
        Character(Len=200) :: line, key, val
        Integer :: u, ios, ieq, ih
        Kl = 0; Kw = 0; KXIJ = 10; MaxNd0 = 3000; kCSF = 0; Ksig = 0; Kdsig = 0; Nconf_neq = 0
        Open(newunit=u, file=path, status='OLD', action='READ', iostat=ios)
        If (ios /= 0) Return
        Do
            Read(u, '(A)', iostat=ios) line
            If (ios /= 0) Exit
            ieq = Index(line, '=')
            If (ieq == 0) Cycle
            If (Index(AdjustL(line), '#') == 1) Cycle
            key = upper(AdjustL(line(1:ieq-1)))
            val = line(ieq+1:)
            ih = Index(val, '#')
            If (ih > 0) val = val(1:ih-1)
            Select Case (Trim(key))
            Case ('KL'); Read(val, *) Kl
            Case ('KW'); Read(val, *) Kw
            Case ('KXIJ'); Read(val, *) KXIJ
            Case ('MAXND0'); Read(val, *) MaxNd0
            Case ('KCSF'); Read(val, *) kCSF
            Case ('KSIG'); Read(val, *) Ksig
            Case ('KDSIG'); Read(val, *) Kdsig
            End Select
        End Do
        Close(u)
        ! MBPT effective integrals only in a Kl = 2 run, which then starts anew
        ! (src/pconf.F90:345-356)
        If (Kl == 2) Then
            Kl = 0
        Else
            Ksig = 0
        End If
    End Subroutine read_ci_in

    ! -------------------------------------------------- CONF.DAT + Init mapping
    Integer Function read_conf_dat(path) Result(rc)
        Character(Len=*), Intent(In) :: path
        Real(dp) :: pq(4*IP6), p(IP6), d
        Integer :: u, ios, lrec, ni, nj, i, nnj, llj, kkj, i0
        rc = 0
        Inquire(iolength=lrec) d
        Open(newunit=u, file=path, status='OLD', access='DIRECT', form='UNFORMATTED', recl=IP6*lrec, iostat=ios)
        If (ios /= 0) Then
            rc = fail('file CONF.DAT is absent')
            Return
        End If
        Read(u, rec=1, iostat=ios) p; pq(1:IP6) = p
        If (ios == 0) Read(u, rec=2, iostat=ios) p; pq(IP6+1:2*IP6) = p
        If (ios == 0) Read(u, rec=5, iostat=ios) p; pq(2*IP6+1:3*IP6) = p
        If (ios == 0) Read(u, rec=6, iostat=ios) p; pq(3*IP6+1:4*IP6) = p
        If (ios == 0 .and. Ksig > 0) Then
            ! HF energies of the orbitals (src/pconf.F90:470, 663-667)
            i0 = Int(pq(3) + 0.01d0)
            Do ni = Nso + 1, Min(Int(pq(2) + 0.01d0), IPS)
                Read(u, rec=2*ni+7, iostat=ios) p
                If (ios /= 0) Exit
                Eps(ni) = -p(i0 + 1)
            End Do
        End If
        Close(u)
        If (ios /= 0) Then
            rc = fail('CONF.DAT: cannot read records 1, 2, 5 and 6 (and 2*ni+7 for Ksig)')
            Return
        End If
        If (Abs(Z - pq(1)) > 1.d-6) Then
            rc = fail('nuclear charge of CONF.DAT differs from CONF.INP')
            Return
        End If
        Ns = Int(pq(2) + 0.01d0)
        If (Ns < 1 .or. Ns > IPS) Then
            rc = fail('CONF.DAT: bad number of orbitals')
            Return
        End If
        Call orbital_table(pq)
        ! map configuration entries to orbitals (src/pconf.F90:506-530)
        Mj = Int(2*Abs(Jm) + 0.01d0)
        If (Allocated(Nip)) Deallocate(Nip, Nq, Nq0, Nvc, Nc0)
        Allocate(Nip(Nsp), Nq(Nsp), Nq0(Nsp), Nvc(Nc), Nc0(Nc))
        Nsu = 0
        Do nj = 1, Nsp
            Call decode_shell_code(Qnl(nj), nnj, llj, kkj, Nq(nj))
            Nip(nj) = orbital_of(nnj, kkj)
            If (Nip(nj) == 0) Then
                rc = fail('no orbital for a shell of CONF.INP')
                Return
            End If
            Nsu = Max(Nsu, Nip(nj))
        End Do
        ! one position per m = -j..j of every orbital up to the highest used
        Nst = Sum(Jj(1:Nsu) + 1)
        rc = split_configurations()
        If (rc /= 0) Return
        Do i = 1, Nsp
            If (Nq(i) > Jj(Nip(i)) + 1) Then
                rc = fail('wrong number of electrons for a shell')
                Return
            End If
        End Do
    End Function read_conf_dat

    ! CONF.INP shell code +-n.lqq (src/pconf.F90:506-516): the digits of |x|
    ! read left to right give the principal number n, the orbital momentum l
    ! and the occupation qq; the sign selects j = l + 1/2 (kappa = -(l+1)) or
    ! j = l - 1/2 (kappa = l).
    Pure Subroutine decode_shell_code(x, n, l, kappa, q)
        Real(dp), Intent(In) :: x
        Integer, Intent(Out) :: n, l, kappa, q
        Real(dp) :: digits
        digits = 10.0d0*(Abs(x) + 1.d-14)
        n = Int(digits)
        digits = 10.0d0*(digits - n)
        l = Int(digits)
        q = Int(100.0d0*(digits - l) + 0.1d0)
        If (x >= 0.0d0) Then
            kappa = -(l + 1)
        Else
            kappa = l
        End If
    End Subroutine decode_shell_code

    ! Index of the basis orbital with principal number n and relativistic
    ! quantum number kappa (0 when CONF.DAT has none).
    Pure Integer Function orbital_of(n, kappa) Result(ni)
        Integer, Intent(In) :: n, kappa
        Do ni = 1, Ns
            If (Nn(ni) == n .and. Kk(ni) == kappa) Return
        End Do
        ni = 0
    End Function orbital_of

    ! Nearest integer of a quantum number stored as a real (half-way values
    ! cannot occur: CONF.DAT holds whole numbers up to rounding).
    Pure Integer Function whole(x)
        Real(dp), Intent(In) :: x
        whole = Int(x + Sign(0.01d0, x))
    End Function whole

    ! The orbital table of CONF.DAT's header records (src/pconf.F90:484-503):
    ! with the marker 0.98765 at word 20, four int32 per orbital (n, l,
    ! kappa, 2j) packed from word 21; otherwise six reals per orbital from
    ! word 21, of which n, l, kappa and 2j are the 1st, 2nd, 5th and 6th.
    Subroutine orbital_table(pq)
        Real(dp), Intent(In) :: pq(:)
        Integer(int32) :: packed(4*IPS)
        Integer :: ni, w
        If (Abs(pq(20) - 0.98765d0) < 1.d-6) Then
            packed = Transfer(pq(21:21+2*IPS-1), packed)
            Do ni = 1, Ns
                Nn(ni) = packed(4*ni-3); Ll(ni) = packed(4*ni-2)
                Kk(ni) = packed(4*ni-1); Jj(ni) = packed(4*ni)
            End Do
        Else
            Do ni = 1, Ns
                w = 20 + 6*(ni - 1)
                Nn(ni) = whole(pq(w+1)); Ll(ni) = whole(pq(w+2))
                Kk(ni) = whole(pq(w+5)); Jj(ni) = whole(pq(w+6))
            End Do
        End If
    End Subroutine orbital_table

    ! Configurations of the entry list (after the Nso core entries): each
    ! runs over consecutive entries until Ne electrons are placed; Nq0 holds
    ! the electrons before an entry within its configuration, Nvc the entries
    ! of a configuration and Nc0 the index before its first entry.
    Integer Function split_configurations() Result(rc)
        Integer :: e, placed, nconf, first, len
        rc = 0
        placed = 0; nconf = 0; first = Nso; len = 0
        Do e = Nso + 1, Nsp
            len = len + 1
            Nq0(e) = placed
            placed = placed + Nq(e)
            If (placed >= Ne) Then
                nconf = nconf + 1
                If (placed > Ne .or. nconf > Nc) Then
                    rc = fail('wrong number of electrons for a configuration')
                    Return
                End If
                Nvc(nconf) = len
                Nc0(nconf) = first
                first = first + len
                placed = 0
                len = 0
            End If
        End Do
        If (nconf /= Nc) rc = fail('configuration list shorter than Nc')
    End Function split_configurations

    ! -------------------------------------------------------------- CONF.INT
    Integer Function read_conf_int(path) Result(rc)
        Character(Len=*), Intent(In) :: path
        Integer :: u, ios, ns1, nso1, nsp1, nsu1, kbrt1, nlist, nrd, i, w
        Integer(int64) :: i8
        Real(real32), Allocatable :: r4(:)
        rc = 0
        Open(newunit=u, file=path, status='OLD', form='UNFORMATTED', action='READ', iostat=ios)
        If (ios /= 0) Then
            rc = fail('Can not find file CONF.INT')
            Return
        End If
        Read(u) ns1, nso1, nsp1, nsu1, Ecore
        If (nso1 /= Nso) Then
            rc = fail('CONF.INT: Nso differs from CONF.INP')
            Close(u)
            Return
        End If
        If (nsu1 < Nsu) Then
            rc = fail('CONF.INT holds fewer orbitals than the configurations use')
            Close(u)
            Return
        End If
        Read(u) (Nn(i), Kk(i), Ll(i), Jj(i), i = 1, Nsu)
        Read(u)
        Read(u) Nhint, kbrt1
        If (kbrt1 /= Kbrt) Then
            rc = fail('Key for Breit of CONF.INT differs from Kbrt of CONF.INP')
            Close(u)
            Return
        End If
        If (Allocated(Rint1)) Deallocate(Rint1, Iint1)
        Allocate(Rint1(Nhint), Iint1(Nhint))
        Read(u) Rint1
        Read(u) Iint1
        Read(u) Ngint, nlist, nrd
        Nx = Int(Sqrt(Real(nrd)) + 0.5)
        w = 1
        If (Kbrt /= 0) w = 2
        If (Allocated(Rint2)) Deallocate(Rint2, Iint2, Iint3)
        Allocate(Rint2(w, Ngint), Iint2(Ngint), Iint3(Ngint))
        ! FP64 integrals (USE_DP_INTEGRALS); an FP32 record is accepted too
        Read(u, iostat=ios) Rint2
        If (ios /= 0) Then
            Backspace(u)
            Allocate(r4(w*Ngint))
            Read(u, iostat=ios) r4
            If (ios /= 0) Then
                rc = fail('CONF.INT: Rint2 record has an unexpected size')
                Close(u)
                Return
            End If
            Do i8 = 1, Ngint
                Rint2(1:w, i8) = Real(r4((i8-1)*w+1:i8*w), dp)
            End Do
        End If
        Read(u) Iint2
        Read(u) Iint3
        Close(u)
    End Function read_conf_int

    ! -------------------------------------------------------------- CONF.GNT
    Integer Function read_conf_gnt(path) Result(rc)
        Character(Len=*), Intent(In) :: path
        Integer :: u, ios
        Integer, Allocatable :: cnt(:)
        rc = 0
        Open(newunit=u, file=path, status='OLD', form='UNFORMATTED', action='READ', iostat=ios)
        If (ios /= 0) Then
            rc = fail('Can not find file CONF.GNT')
            Return
        End If
        Read(u, iostat=ios) Nlx, Ngaunt
        If (ios /= 0) Then
            Rewind(u)
            Read(u) Ngaunt
            Nlx = 5
        Else
            Allocate(cnt(Nlx + 1))
            Read(u) cnt
        End If
        If (Allocated(Gin)) Deallocate(Gin, Gnt)
        Allocate(Gin(Ngaunt), Gnt(Ngaunt))
        Read(u) Gin
        Read(u) Gnt
        Close(u)
    End Function read_conf_gnt

    ! ------------------------------------------------------- Dinit + Det_List
    Subroutine dinit
        Integer :: ni, j, i0
        i0 = 0
        Do ni = 1, Ns
            i0 = i0 + Jj(ni) + 1
        End Do
        Npos = i0
        If (Allocated(Jz)) Deallocate(Jz, Nh)
        Allocate(Jz(Npos), Nh(Npos))
        i0 = 0
        Do ni = 1, Ns
            Nf0(ni) = i0
            Do j = 1, 2*Jj(ni) + 1, 2
                i0 = i0 + 1
                Jz(i0) = j - (Jj(ni) + 1)
                Nh(i0) = ni
            End Do
        End Do
    End Subroutine dinit

    ! Next determinant of configuration ic in odometer order: the last entry
    ! of the configuration varies fastest, inside an entry the last electron.
    Subroutine ndet(ic, fin, idet, msum)
        Integer, Intent(In) :: ic
        Logical, Intent(InOut) :: fin
        Integer, Intent(InOut) :: idet(:)
        Integer, Intent(Out) :: msum
        Integer :: n1, n2, ni, nqi, i, n0, iq, is, im, top, k, jq, nj
        n1 = Nc0(ic) + 1
        n2 = Nc0(ic) + Nvc(ic)
        msum = 0
        If (fin) Then
            Do ni = n1, n2
                nqi = Nq(ni)
                i = Nip(ni)
                n0 = Nq0(ni)
                Do iq = n0 + 1, n0 + nqi
                    idet(iq) = Nf0(i) + iq - n0
                End Do
            End Do
            fin = .false.
        Else
            Do is = n2, n1, -1
                nqi = Nq(is)
                If (nqi == 0) Cycle
                i = Nip(is)
                n0 = Nq0(is)
                im = n0 + nqi
                top = Nf0(i) + Jj(i) + 1       ! last position of the shell
                Do k = im, n0 + 1, -1
                    If (idet(k) < top - (im - k)) Then
                        idet(k) = idet(k) + 1
                        Do jq = k + 1, im
                            idet(jq) = idet(k) + (jq - k)
                        End Do
                        Do nj = is + 1, n2
                            n0 = Nq0(nj)
                            Do jq = n0 + 1, n0 + Nq(nj)
                                idet(jq) = Nf0(Nip(nj)) + jq - n0
                            End Do
                        End Do
                        Go To 10
                    End If
                End Do
            End Do
            fin = .true.
            Return
        End If
10      Do iq = 1, Ne
            msum = msum + Jz(idet(iq))
        End Do
    End Subroutine ndet

    Subroutine enumerate_dets
        Integer :: ic, msum, ndi, pass, a, b, k, t
        Integer :: idet(Ne), sdet(Ne)
        Logical :: fin
        If (Allocated(Ndc)) Deallocate(Ndc)
        Allocate(Ndc(Nc))
        Do pass = 1, 2
            Nd = 0
            Do ic = 1, Nc
                fin = .true.
                ndi = 0
                Call ndet(ic, fin, idet, msum)
                Do While (.not. fin)
                    If (msum == Mj) Then
                        ndi = ndi + 1
                        If (pass == 2) Then
                            sdet = idet
                            Do a = 1, Ne - 1
                                k = a
                                Do b = a + 1, Ne
                                    If (sdet(b) < sdet(k)) k = b
                                End Do
                                If (k /= a) Then
                                    t = sdet(k); sdet(k) = sdet(a); sdet(a) = t
                                End If
                            End Do
                            Iarr(:, Nd + ndi) = sdet
                        End If
                    End If
                    Call ndet(ic, fin, idet, msum)
                End Do
                Ndc(ic) = ndi
                Nd = Nd + ndi
            End Do
            If (pass == 1) Then
                If (Allocated(Iarr)) Deallocate(Iarr)
                Allocate(Iarr(Ne, Max(Nd, 1)))
            End If
        End Do
    End Subroutine enumerate_dets

    ! J terms of the space (Jterm, src/determinants.f90:185-348): for each
    ! configuration with determinants at M = Mj, count its determinants per
    ! M >= Mj; walking down from the top M, what is left at each M is the
    ! number of new levels with J = M.  jt: doubled J values, ascending.
    Integer Function jterms(jt, cap) Result(njd)
        Integer, Intent(In) :: cap
        Integer, Intent(Out) :: jt(cap)
        Integer :: ic, msum, ndi, im, imax, jmax, mt, j, n, t
        Integer :: idet(Ne)
        Integer, Allocatable :: nmj(:)
        Logical :: fin, swapped
        njd = 0
        jmax = 1
        Do ic = 1, Nc
            fin = .true.
            Call ndet(ic, fin, idet, msum)
            Do While (.not. fin)
                jmax = Max(jmax, (msum - Mj) / 2 + 1)
                Call ndet(ic, fin, idet, msum)
            End Do
        End Do
        Allocate(nmj(jmax))
        Do ic = 1, Nc
            fin = .true.
            ndi = 0
            imax = 1
            Call ndet(ic, fin, idet, msum)
            Do While (.not. fin)
                If (msum >= Mj) Then
                    If (msum == Mj) ndi = ndi + 1
                    imax = Max(imax, (msum - Mj) / 2 + 1)
                End If
                Call ndet(ic, fin, idet, msum)
            End Do
            If (ndi == 0) Cycle
            nmj(1:imax) = 0
            fin = .true.
            Call ndet(ic, fin, idet, msum)
            Do While (.not. fin)
                If (msum >= Mj) nmj((msum - Mj) / 2 + 1) = nmj((msum - Mj) / 2 + 1) + 1
                Call ndet(ic, fin, idet, msum)
            End Do
            Do im = imax, 1, -1
                n = nmj(im)
                If (n == 0) Cycle
                mt = (im - 1) * 2 + Mj
                nmj(1:im) = nmj(1:im) - n
                Do j = 1, njd
                    If (jt(j) == mt) Exit
                End Do
                If (j > njd) Then
                    If (njd >= cap) Then
                        njd = -1
                        Return
                    End If
                    njd = njd + 1
                    jt(njd) = mt
                End If
            End Do
        End Do
        Do
            swapped = .false.
            Do j = 2, njd
                If (jt(j) < jt(j - 1)) Then
                    t = jt(j); jt(j) = jt(j - 1); jt(j - 1) = t
                    swapped = .true.
                End If
            End Do
            If (.not. swapped) Exit
        End Do
        Deallocate(nmj)
    End Function jterms

    Integer(c_int) Function pconf_jterms(jt, cap) bind(C, name="pconf_jterms") Result(njd)
        Integer(c_int), Value :: cap
        Integer(c_int), Intent(Out) :: jt(cap)
        njd = jterms(jt, cap)
    End Function pconf_jterms

    ! Nd0: largest prefix of the first Nc4 configurations holding <= MaxNd0
    ! basis functions (src/determinants.f90:22-45): determinants, or with
    ! ndcs (CSFs per configuration class) given, CSFs
    Subroutine calc_nd0(ndcs)
        Integer, Intent(In), Optional :: ndcs(:)
        Integer :: ic, n1
        Nc1 = 0
        Nd0 = 0
        n1 = 0
        Do ic = 1, Min(Nc4, Nc)
            If (Present(ndcs)) Then
                n1 = n1 + ndcs(Nc_neq(ic))
            Else
                n1 = n1 + Ndc(ic)
            End If
            If (n1 < MaxNd0 + 1) Then
                Nd0 = n1
                Nc1 = ic
            End If
        End Do
    End Subroutine calc_nd0

    ! ------------------------------------------------------ nonequiv_conf
    ! Classes of equivalent configurations for the CSF basis
    ! (src/csf.f90:454-536), formed before the determinants are listed.  A
    ! configuration joins the class of the first earlier one whose entries
    ! its own match one to one in l, 2j and occupation (each of its entries
    ! takes the first unclaimed matching entry of the other); its entries are
    ! then reordered so that they line up with the other's, and the two list
    ! their determinants in corresponding order -- one set of CSF
    ! coefficients serves the class.  The reordering is pconf's: for each
    ! entry position p of the other configuration in turn, every entry of
    ! this one that matches the other's entry p is swapped into position p
    ! (pconf keeps a claim mark per shell through the swaps; for this
    ! configuration's shells the marks are clear and stay so).
    Subroutine nonequiv_conf
        Integer :: ic, jc, i0, i1, j0, j1, nii, njj, nii1, p, q, t
        Integer :: mark(IPS)
        Logical :: hit
        Write(*, '(/4x,a)') 'Creating list of nonequivalent  configurations...'
        If (Allocated(Nc_neq)) Deallocate(Nc_neq)
        Allocate(Nc_neq(Nc))
        Nconf_neq = 0
        mark = 0
        conf: Do ic = 1, Nc
            i0 = Nc0(ic) + 1
            i1 = Nc0(ic) + Nvc(ic)
            other: Do jc = 1, ic - 1
                j0 = Nc0(jc) + 1
                j1 = Nc0(jc) + Nvc(jc)
                mark(Nip(j0:j1)) = 0
                Do nii = i0, i1
                    hit = .false.
                    Do njj = j0, j1
                        If (mark(Nip(njj)) == 1) Cycle
                        If (.not. alike(nii, njj)) Cycle
                        mark(Nip(njj)) = 1
                        hit = .true.
                        Exit
                    End Do
                    If (.not. hit) Cycle other
                End Do
                mark(Nip(i0:i1)) = 0
                Do p = 1, j1 - j0 + 1
                    njj = j0 + p - 1
                    Do q = 1, i1 - i0 + 1
                        nii = i0 + q - 1
                        If (mark(Nip(nii)) == 1) Cycle
                        If (.not. alike(nii, njj)) Cycle
                        If (q /= p) Then
                            nii1 = i0 + p - 1
                            t = mark(Nip(nii1)); mark(Nip(nii1)) = mark(Nip(nii)); mark(Nip(nii)) = t
                            t = Nip(nii1); Nip(nii1) = Nip(nii); Nip(nii) = t
                            t = Nq(nii1); Nq(nii1) = Nq(nii); Nq(nii) = t
                        End If
                    End Do
                End Do
                Nc_neq(ic) = Nc_neq(jc)
                Cycle conf
            End Do other
            Nconf_neq = Nconf_neq + 1
            Nc_neq(ic) = Nconf_neq
        End Do conf
        ! electrons before each entry within its configuration, over the new order
        Do ic = 1, Nc
            t = 0
            Do nii = Nc0(ic) + 1, Nc0(ic) + Nvc(ic)
                Nq0(nii) = t
                t = t + Nq(nii)
            End Do
        End Do
        Write(*, '(4x,"Number of nonequivalent configurations:",i7,/)') Nconf_neq
    Contains
        Logical Function alike(a, b)
            Integer, Intent(In) :: a, b
            alike = Ll(Nip(a)) == Ll(Nip(b)) .and. Jj(Nip(a)) == Jj(Nip(b)) .and. Nq(a) == Nq(b)
        End Function alike
    End Subroutine nonequiv_conf

    ! Full start-up in directory dir (no trailing slash needed).
    Integer Function load_space(dir) Result(rc)
        Character(Len=*), Intent(In) :: dir
        Character(Len=1024) :: p
        p = Trim(dir)//'/'
        rc = read_conf_inp(Trim(p)//'CONF.INP'); If (rc /= 0) Return
        Call read_ci_in(Trim(p)//'ci.in')
        rc = read_conf_dat(Trim(p)//'CONF.DAT'); If (rc /= 0) Return
        rc = read_conf_int(Trim(p)//'CONF.INT'); If (rc /= 0) Return
        rc = read_conf_gnt(Trim(p)//'CONF.GNT'); If (rc /= 0) Return
        ! CSF basis: configuration classes before Dinit (src/pconf.F90:138)
        If (kCSF > 0) Call nonequiv_conf
        Call dinit
        Call enumerate_dets
        If (Nd == 0) Then
            rc = fail('term J is absent in all configurations')
            Return
        End If
        Call calc_nd0
        ! FormB0 takes at most Nd0 eigenvectors of the initial block
        If (Nd0 > 0 .and. Nlv > Nd0) Nlv = Nd0
        If (Ksig > 0) Then
            rc = read_sgc_con(Trim(p)//'SGC.CON'); If (rc /= 0) Return
            If (Ksig == 2) Then
                rc = read_scrc_con(Trim(p)//'SCRC.CON'); If (rc /= 0) Return
            End If
            If (Kdsig /= 0) Then
                If (.not. have_e0) Then
                    ! pconf asks for it on the terminal (src/pconf.F90:697-704)
                    Write(*, '(5X,"Give valence energy: ",$)')
                    Read(*, *, iostat=rc) E_0
                    If (rc /= 0) Then
                        rc = fail('Kdsig = 1 needs the valence energy E_0 on standard input')
                        Return
                    End If
                    have_e0 = .true.
                End If
                Call form_d
            End If
        End If
    End Function load_space

    ! RintS, SGC.CON part (src/pconf.F90:715-775): header (7X,I3,5X,I1,24X,I1),
    ! then one list-directed line per pair na <= nb with equal l and j, in
    ! pconf's loop order, values read in single precision; with Ksig = 2 a
    ! title line and the ten screening coefficients as (5X,I2,5X,F8.5).
    Integer Function read_sgc_con(path) Result(rc)
        Character(Len=*), Intent(In) :: path
        Integer :: u, ios, na, nb, na1, nb1, khot, i, pass, k, k1
        Character(Len=4) :: idummy
        Real(real32) :: x, y, zz
        rc = 0
        Open(newunit=u, file=path, status='OLD', action='READ', iostat=ios)
        If (ios /= 0) Then
            rc = fail('file SGC.CON is absent')
            Return
        End If
        Do pass = 1, 2
            Rewind(u)
            Read(u, '(7X,I3,5X,I1,24X,I1)', iostat=ios) NmaxS, LmaxS, khot
            If (ios /= 0 .or. NmaxS > Ns) Then
                rc = fail('SGC.CON: bad header')
                Close(u)
                Return
            End If
            i = 0
            Do na = Nso + 1, NmaxS
                If (Ll(na) > LmaxS) Cycle
                Do nb = na, NmaxS
                    If (Ll(nb) /= Ll(na) .or. Jj(nb) /= Jj(na)) Cycle
                    i = i + 1
                    If (pass == 1) Cycle
                    Read(u, *, iostat=ios) idummy, na1, nb1, x, y, zz
                    If (ios /= 0 .or. na1 /= na .or. nb1 /= nb) Then
                        rc = fail('SGC.CON: wrong indices for a sigma pair')
                        Close(u)
                        Return
                    End If
                    Iint1S(i) = Nx*(na - Nso - 1) + (nb - Nso)
                    Rsig(i) = x
                    Dsig(i) = y
                    Esig(i) = zz
                End Do
            End Do
            If (pass == 1) Then
                NhintS = i
                If (Allocated(Iint1S)) Deallocate(Iint1S, Rsig, Dsig, Esig)
                Allocate(Iint1S(NhintS), Rsig(NhintS), Dsig(NhintS), Esig(NhintS))
            End If
        End Do
        If (Ksig == 2) Then
            Read(u, '(5X,I2,5X,F8.5)', iostat=ios)
            Do k = 1, 10
                If (ios == 0) Read(u, '(5X,I2,5X,F8.5)', iostat=ios) k1, Scr(k)
                If (ios /= 0 .or. k1 /= k - 1) Then
                    rc = fail('SGC.CON: wrong order of screening coefficients')
                    Close(u)
                    Return
                End If
            End Do
        End If
        Close(u)
    End Function read_sgc_con

    ! RintS, SCRC.CON part (src/pconf.F90:781-840): the two-electron sigma
    ! integrals of Ksig = 2 with their keys, Ksym, Kmax and Nsum (4 Nav, or
    ! 4 NmaxS when Nsx2 != 0); the radial values in FP64 (USE_DP_INTEGRALS).
    Integer Function read_scrc_con(path) Result(rc)
        Character(Len=*), Intent(In) :: path
        Integer :: u, ios, ns1, nso1, nsh1, khot1, nsx1, nsx2, nmax1, lmax1, nav, kbox, nrd, i
        Integer, Allocatable :: ord(:)
        rc = 0
        Open(newunit=u, file=path, status='OLD', form='UNFORMATTED', action='READ', iostat=ios)
        If (ios /= 0) Then
            rc = fail('file SCRC.CON is absent')
            Return
        End If
        Read(u, iostat=ios) ns1, nso1, nsh1, khot1, nsx1, nsx2, Ksym
        If (ios == 0) Read(u, iostat=ios) nmax1, lmax1, KmaxS, nav, kbox
        If (ios == 0) Read(u, iostat=ios) NgintS, nrd
        If (ios /= 0) Then
            rc = fail('SCRC.CON: cannot read its header records')
            Close(u)
            Return
        End If
        If (nso1 /= Nso .or. nrd /= Nx*Nx .or. NmaxS < nmax1 .or. LmaxS < lmax1) Then
            rc = fail('SCRC.CON does not match CONF.INT / SGC.CON (Nso, IPx, NmaxS or LmaxS)')
            Close(u)
            Return
        End If
        If (nsx2 == 0) Then
            Nsum = 4*nav
        Else
            Nsum = 4*nmax1
        End If
        If (Allocated(Iint2S)) Deallocate(Iint2S, Iint3S, Rint2S, Dint2S, Eint2S)
        Allocate(Iint2S(NgintS), Iint3S(NgintS), Rint2S(NgintS), Dint2S(NgintS), Eint2S(NgintS), ord(nrd))
        Read(u, iostat=ios) (Rint2S(i), i = 1, NgintS)
        If (ios == 0) Read(u, iostat=ios) (Iint2S(i), i = 1, NgintS)
        If (ios == 0) Read(u, iostat=ios) (Iint3S(i), i = 1, NgintS)
        If (ios == 0) Read(u, iostat=ios) (ord(i), i = 1, nrd)
        If (ios == 0) Read(u, iostat=ios) (Dint2S(i), i = 1, NgintS)
        If (ios == 0) Read(u, iostat=ios) (Eint2S(i), i = 1, NgintS)
        Close(u)
        If (ios /= 0) rc = fail('SCRC.CON: a table record is short')
    End Function read_scrc_con

    ! FormD (src/pconf.F90:852-880): E_0 minus the HF energies of each
    ! determinant's orbitals, the E_k of the energy-dependent sigma
    Subroutine form_d
        Integer :: n, i
        Real(dp) :: x
        If (Allocated(DiagS)) Deallocate(DiagS)
        Allocate(DiagS(Nd))
        Do n = 1, Nd
            x = 0.0_dp
            Do i = 1, Ne
                x = x + Eps(Nh(Iarr(i, n)))
            End Do
            DiagS(n) = E_0 - x
        End Do
    End Subroutine form_d

    ! Space and integrals to the device context.
    Integer Function upload(ctx) Result(rc)
        Type(c_ptr), Intent(In) :: ctx
        Integer(c_int32_t), Allocatable :: ll32(:), jj32(:), kk32(:)
        Real(dp), Allocatable :: r2(:)
        Allocate(ll32(Ns), jj32(Ns), kk32(Ns))
        ll32 = Ll(1:Ns); jj32 = Jj(1:Ns); kk32 = Kk(1:Ns)
        rc = pci_gpu_set_space(ctx, Ns, Nso, ll32, jj32, kk32, Npos, Jz, Nh, Ne, Nd, Iarr, Nc, Ndc, Mj, Gj)
        If (rc /= 0) Then
            host_err = pci_error_text()
            Return
        End If
        Allocate(r2(Size(Rint2)))
        r2 = Reshape(Rint2, [Size(Rint2)])
        rc = pci_gpu_set_integrals(ctx, Nx, Nhint, Rint1, Iint1, Ngint, Kbrt, r2, Iint2, Iint3, Nlx, Ngaunt, Gin, Gnt)
        If (rc /= 0) Then
            host_err = pci_error_text()
            Return
        End If
        If (Ksig > 0) Then
            If (Kdsig /= 0) Then
                rc = pci_gpu_set_sigma(ctx, Ksig, Kdsig, NmaxS, LmaxS, NhintS, Iint1S, Rsig, Dsig, Esig, Eps(1:Ns), &
                                       c_loc(DiagS))
            Else
                rc = pci_gpu_set_sigma(ctx, Ksig, Kdsig, NmaxS, LmaxS, NhintS, Iint1S, Rsig, Dsig, Esig, Eps(1:Ns), &
                                       c_null_ptr)
            End If
            If (rc /= 0) Then
                host_err = pci_error_text()
                Return
            End If
            If (Ksig == 2) Then
                rc = pci_gpu_set_sigma2(ctx, Ksym, KmaxS, Nsum, Scr, Int(NgintS, c_int64_t), Iint2S, Iint3S, &
                                        Rint2S, Dint2S, Eint2S)
                If (rc /= 0) host_err = pci_error_text()
            End If
        End If
    End Function upload

    ! -------------------------------------------------------------- outputs
    Subroutine write_conf_det(path)
        Character(Len=*), Intent(In) :: path
        Integer :: u, i
        Open(newunit=u, file=path, status='UNKNOWN', form='UNFORMATTED')
        Write(u) Nd, Nsu
        Do i = 1, Nd
            Write(u) Iarr(1:Ne, i)
        End Do
        Close(u)
    End Subroutine write_conf_det

    Subroutine write_hij(path, rowptr, col, val)
        Character(Len=*), Intent(In) :: path
        Integer(int64), Intent(In) :: rowptr(0:)
        Integer(int32), Intent(In) :: col(:)
        Real(dp), Intent(In) :: val(:)
        Integer :: u
        Integer(int64) :: nd8, nnz
        nd8 = Size(rowptr) - 1
        nnz = rowptr(nd8)
        Open(newunit=u, file=path, status='REPLACE', access='STREAM', form='UNFORMATTED')
        Write(u) nd8, nnz, rowptr, col(1:nnz), val(1:nnz)
        Close(u)
    End Subroutine write_hij

    ! Legacy triple file CONF.HIJ / CONF.JJJ, as the sort program writes it
    ! for pol and ine (src/sort.f90:64-70, read by src/pol.f90:766-798): an
    ! int64 count, then (row, column, value) records as int32, int32, real64,
    ! ordered by row then column, exact zeros left out.  The CSR rows hold
    ! ascending columns, so row order is that order.
    Subroutine write_hij_triples(path, rowptr, col, val)
        Character(Len=*), Intent(In) :: path
        Integer(int64), Intent(In) :: rowptr(0:)
        Integer(int32), Intent(In) :: col(:)
        Real(dp), Intent(In) :: val(:)
        Type, Bind(C) :: hij_rec
            Integer(c_int32_t) :: n, k
            Real(c_double) :: t
        End Type hij_rec
        Integer, Parameter :: CHUNK = 65536
        Type(hij_rec), Allocatable :: buf(:)
        Integer :: u, m
        Integer(int64) :: nd8, n, l, cnt
        nd8 = Size(rowptr) - 1
        cnt = Count(val(1:rowptr(nd8)) /= 0.0_dp)
        Allocate(buf(CHUNK))
        Open(newunit=u, file=path, status='REPLACE', access='STREAM', form='UNFORMATTED')
        Write(u) cnt
        m = 0
        Do n = 1, nd8
            Do l = rowptr(n - 1) + 1, rowptr(n)
                If (val(l) == 0.0_dp) Cycle
                m = m + 1
                buf(m) = hij_rec(Int(n, c_int32_t), col(l), val(l))
                If (m == CHUNK) Then
                    Write(u) buf(1:m)
                    m = 0
                End If
            End Do
        End Do
        If (m > 0) Write(u) buf(1:m)
        Close(u)
        Deallocate(buf)
    End Subroutine write_hij_triples

    Integer Function read_hij(path, rowptr, col, val) Result(rc)
        Character(Len=*), Intent(In) :: path
        Integer(int64), Allocatable, Intent(Out) :: rowptr(:)
        Integer(int32), Allocatable, Intent(Out) :: col(:)
        Real(dp), Allocatable, Intent(Out) :: val(:)
        Integer :: u, ios
        Integer(int64) :: nd8, nnz
        rc = 0
        Open(newunit=u, file=path, status='OLD', access='STREAM', form='UNFORMATTED', action='READ', iostat=ios)
        If (ios /= 0) Then
            rc = fail('cannot open '//Trim(path))
            Return
        End If
        Read(u) nd8, nnz
        If (nd8 /= Nd) Then
            rc = fail(Trim(path)//' was written for another determinant space')
            Close(u)
            Return
        End If
        Allocate(rowptr(0:nd8), col(nnz), val(nnz))
        Read(u) rowptr, col, val
        Close(u)
    End Function read_hij

    Subroutine write_xij(path, tk, tj, vecs)
        Character(Len=*), Intent(In) :: path
        Real(dp), Intent(In) :: tk(:), tj(:), vecs(:, :)   ! vecs(basis size, Nlv)
        Integer :: u, n, nb
        nb = Size(vecs, 1)
        Open(newunit=u, file=path, status='UNKNOWN', form='UNFORMATTED')
        Do n = 1, Size(tk)
            Write(u) tk(n), tj(n), nb, vecs(:, n)
        End Do
        Close(u)
    End Subroutine write_xij

    ! jbasis's CSF.LOG (src/csf.f90:88-95, 202): one line per class in the
    ! order jbasis takes them, by descending determinant count (ties in class
    ! order): class, its first configuration, determinants, CSFs, core (0).
    Subroutine write_csf_log(ndcs)
        Integer, Intent(In) :: ndcs(:)
        Integer :: first(Nconf_neq), ndq(Nconf_neq), order(Nconf_neq)
        Integer :: u, ic, q, i, j, k
        first = 0
        ndq = 0
        Do ic = 1, Nc
            q = Nc_neq(ic)
            If (first(q) == 0) Then
                first(q) = ic
                ndq(q) = Ndc(ic)
            End If
        End Do
        Do i = 1, Nconf_neq
            order(i) = i
        End Do
        Do i = 2, Nconf_neq  ! stable insertion sort, descending
            j = order(i)
            k = i - 1
            Do While (k >= 1)
                If (ndq(order(k)) >= ndq(j)) Exit
                order(k + 1) = order(k)
                k = k - 1
            End Do
            order(k + 1) = j
        End Do
        Open(newunit=u, file='CSF.LOG', status='UNKNOWN')
        Write(u, '(6x,"Iconf_neq",3x,"Iconf",4x,"Ndi",5x,"Ncsfi",5x,"Core")')
        Do i = 1, Nconf_neq
            q = order(i)
            Write(u, '(6x,i5,3x,i7,4x,i5,3(4x,i5))') q, first(q), ndq(q), ndcs(q), 0
        End Do
        Close(u)
    End Subroutine write_csf_log

    ! FormB0's |Kl4| = 2 start (src/davidson.f90:136-154): up to nmax
    ! records of CONF.XIJ (CONF.WFS in the CSF basis) -- energy, J, length,
    ! vector -- as start energies tk0 and vectors x0(nbas, n), the vector
    ! length taken from the first record and zero-padded to nbas.  n = the
    ! records read (pconf's Nlv = num).
    Integer Function read_start_vectors(path, nbas, nmax, tk0, x0, n, ndpt) Result(rc)
        Character(Len=*), Intent(In) :: path
        Integer, Intent(In) :: nbas, nmax
        Real(dp), Allocatable, Intent(Out) :: tk0(:), x0(:, :)
        Integer, Intent(Out) :: n, ndpt
        Integer :: u, ios, j, idum
        Real(dp) :: e, t
        rc = 0
        n = 0
        ndpt = 0
        Open(newunit=u, file=path, status='OLD', form='UNFORMATTED', action='READ', iostat=ios)
        If (ios /= 0) Then
            rc = fail('Kl4 = 2: '//path//' is absent')
            Return
        End If
        Read(u, iostat=ios) e, t, ndpt
        If (ios /= 0 .or. ndpt < 1 .or. ndpt > nbas) Then
            Close(u)
            rc = fail('Kl4 = 2: no vectors of this basis in '//path)
            Return
        End If
        Rewind(u)
        Allocate(tk0(nmax), x0(nbas, nmax))
        tk0 = 0
        x0 = 0
        Do j = 1, nmax
            Read(u, iostat=ios) tk0(j), t, idum, x0(1:ndpt, j)
            If (ios /= 0) Exit
            n = j
        End Do
        Close(u)
    End Function read_start_vectors

    ! The KXIJ checkpoint (FormB, src/davidson.f90:219-263): the current
    ! vectors to CONF.XIJ, or CONF.WFS in the CSF basis with J = Jm
    Subroutine xij_checkpoint(user, it, nlv_c, nd_c, tk, tj, vecs) bind(C)
        Type(c_ptr), Value :: user
        Integer(c_int32_t), Value :: it, nlv_c, nd_c
        Real(c_double), Intent(In) :: tk(nlv_c), tj(nlv_c), vecs(nd_c, nlv_c)
        Real(dp) :: tjw(nlv_c)
        If (kCSF > 0) Then
            tjw = Jm
            Call write_xij('CONF.WFS', tk, tjw, vecs)
        Else
            tjw = tj
            Call write_xij('CONF.XIJ', tk, tj, vecs)
        End If
        If (it > 0) Then
            Write(*, '(A,I0)') ' FormB: Vectors saved at iteration ', it
            Call write_conf_eng(it, tk, tjw, nd_c)
        End If
    End Subroutine xij_checkpoint

    ! PrintEnergiesDvdsn (src/pconf.F90:2266-2345): the energy table of
    ! iteration it to CONF.ENG, a new file at the first checkpoint and
    ! appended (after a blank line) at the later ones.  nbas: basis size.
    Subroutine write_conf_eng(it, tk, tj, nbas)
        Integer, Intent(In) :: it, nbas
        Real(dp), Intent(In) :: tk(:), tj(:)
        Character(Len=7), Parameter :: kind(9) = ['COULOMB', 'C+MBPT1', 'C+MBPT2', 'GAUNT  ', 'G+MBPT1', &
                                                  'G+MBPT2', 'BREIT  ', 'B+MBPT1', 'B+MBPT2']
        Real(dp) :: t0(Size(tk))
        Integer :: u, j, ist
        If (it <= KXIJ) Then
            Open(newunit=u, file='CONF.ENG', status='REPLACE', action='WRITE')
        Else
            Open(newunit=u, file='CONF.ENG', status='UNKNOWN', position='APPEND', action='WRITE')
            Write(u, '(A)') ''
        End If
        ist = Min(Ksig, 2) + 1 + 3*Min(Kbrt, 2)
        Write(u, '(A,I3)') 'Davidson Iteration #', it
        Write(u, '(4X,63("="))')
        If (Ksig*Kdsig == 0) Then
            Write(u, '(4X,"Energy levels (",A7," Nc=",I7," Nd=",I9,"); Gj =",F7.4, &
                     &/4X,"N",6X,"JTOT",12X,"EV",16X,"ET",9X,"DEL(CM**-1)")') kind(ist), Nc, nbas, Gj
        Else
            Write(u, '(4X,"Energy levels ",A7,", Sigma(E =",F10.4,") extrapolation var.",I2, &
                     &/4X,"(Nc=",I7," Nd=",I9,"); Gj =",F7.4,/4X,"N",6X,"JTOT",12X,"EV",16X,"ET",9X, &
                     &"DEL(CM**-1)")') kind(ist), E_0, 1, Nc, nbas, Gj
        End If
        Write(u, '(4X,63("-"))')
        Do j = 1, Size(tk)
            t0(j) = tk(j) + 4.d0*Gj*tj(j)*(tj(j) + 1.d0)
        End Do
        Do j = 1, Size(tk)
            Write(u, '(3X,I3,F14.9,F14.8,F18.6,F15.2)') j, tj(j), t0(j), t0(j) - Ecore, (t0(1) - t0(j))*2*DPRy
        End Do
        Write(u, '(4X,63("="))')
        Close(u)
    End Subroutine write_conf_eng

    ! Energy table of PrintEnergies (src/pconf.F90:3152-3242); ncsf present:
    ! the header of a CSF-basis run
    Subroutine print_energies(unit, tk_in, tj, ncsf)
        Integer, Intent(In) :: unit
        Real(dp), Intent(In) :: tk_in(:), tj(:)
        Integer, Intent(In), Optional :: ncsf
        Character(Len=7), Parameter :: interaction(3) = ['COULOMB', 'GAUNT  ', 'BREIT  ']
        Character(Len=3), Parameter :: effective(3) = [' NO', ' 1e', 'YES']
        Real(dp) :: tk(Size(tk_in)), dt, del
        Integer :: j
        Write(unit, '(1X,80("="))')
        If (Present(ncsf)) Then
            Write(unit, '(" Energy levels for ",A16," (Nc=",I7," Nd=",I9," nbas=",I9,")",/" Interaction: ",A7, &
                         &"; Effective integrals: ",A3,"; Symmetry: ",A3,/"  N",9X,"JTOT",12X,"EV",15X,"ET",8X, &
                         &"DEL(CM**-1)")') confname, Nc, Nd, ncsf, interaction(Min(Kbrt, 2) + 1), &
                         effective(Min(Ksig, 2) + 1), 'CSF'
        Else
            Write(unit, '(4X,"Energy levels (",A7," Nc=",I7," Nd=",I9,"); Gj =",F7.4, &
                         &/4X,"N",6X,"JTOT",12X,"EV",16X,"ET",9X,"DEL(CM**-1)")') 'COULOMB', Nc, Nd, Gj
        End If
        Write(unit, '(1X,80("-"))')
        Do j = 1, Size(tk_in)
            tk(j) = tk_in(j) + 4.d0*Gj*tj(j)*(tj(j) + 1.d0)
        End Do
        Do j = 1, Size(tk)
            dt = tk(j) - Ecore
            del = (tk(1) - tk(j))*2*DPRy
            Write(unit, '(3X,I3,F14.9,F14.8,F18.6,F15.2)') j, tj(j), tk(j), dt, del
        End Do
        Write(unit, '(1X,80("="))')
    End Subroutine print_energies

    ! ------------------------------------------------- bind(C) entry points
    Function c_to_f(s) Result(f)
        Character(Kind=c_char), Intent(In) :: s(*)
        Character(Len=1024) :: f
        Integer :: i
        f = ' '
        Do i = 1, 1024
            If (s(i) == c_null_char) Exit
            f(i:i) = s(i)
        End Do
    End Function c_to_f

    Integer(c_int) Function pconf_load(dir) bind(C, name="pconf_load")
        Character(Kind=c_char), Intent(In) :: dir(*)
        pconf_load = load_space(Trim(c_to_f(dir)))
    End Function pconf_load

    ! dims: Nd Ne Nc Nst Npos Ns Nso Nsu Nlv Nd0 N_it Kl Kw Mj Kbrt Ngint Nhint Ngaunt KXIJ Kv Nc1 Kl4
    Subroutine pconf_dims(d) bind(C, name="pconf_dims")
        Integer(c_int64_t), Intent(Out) :: d(24)
        d = 0
        d(1:22) = [Int(Nd, int64), Int(Ne, int64), Int(Nc, int64), Int(Nst, int64), Int(Npos, int64), &
                   Int(Ns, int64), Int(Nso, int64), Int(Nsu, int64), Int(Nlv, int64), Int(Nd0, int64), &
                   Int(N_it, int64), Int(Kl, int64), Int(Kw, int64), Int(Mj, int64), Int(Kbrt, int64), Ngint, &
                   Int(Nhint, int64), Int(Ngaunt, int64), Int(KXIJ, int64), Int(Kv, int64), Int(Nc1, int64), &
                   Int(Kl4, int64)]
    End Subroutine pconf_dims

    ! params: Crt4 Gj Z Jm Ecore XJ_av
    Subroutine pconf_params(r) bind(C, name="pconf_params")
        Real(c_double), Intent(Out) :: r(8)
        r = 0
        r(1:6) = [Crt4, Gj, Z, Jm, Ecore, XJ_av]
    End Subroutine pconf_params

    Subroutine pconf_space(iarr_o, ndc_o, jz_o, nh_o, ll_o, jj_o, kk_o) bind(C, name="pconf_space")
        Integer(c_int32_t), Intent(Out) :: iarr_o(Ne, Nd), ndc_o(Nc), jz_o(Npos), nh_o(Npos), ll_o(Ns), jj_o(Ns), &
                                           kk_o(Ns)
        iarr_o = Iarr(:, 1:Nd)
        ndc_o = Ndc
        jz_o = Jz
        nh_o = Nh
        ll_o = Ll(1:Ns)
        jj_o = Jj(1:Ns)
        kk_o = Kk(1:Ns)
    End Subroutine pconf_space

    Integer(c_int) Function pconf_upload(ctx) bind(C, name="pconf_upload")
        Type(c_ptr), Value :: ctx
        pconf_upload = upload(ctx)
    End Function pconf_upload

    ! CONF.HIJ-style triple file from a 1-based lower-triangle CSR
    Subroutine pconf_write_triples(path, nd, rowptr, col, val) bind(C, name="pconf_write_triples")
        Character(Kind=c_char), Intent(In) :: path(*)
        Integer(c_int), Value :: nd
        Integer(c_int64_t), Intent(In) :: rowptr(0:nd)
        Integer(c_int32_t), Intent(In) :: col(*)
        Real(c_double), Intent(In) :: val(*)
        Call write_hij_triples(Trim(c_to_f(path)), rowptr, col(1:rowptr(nd)), val(1:rowptr(nd)))
    End Subroutine pconf_write_triples

    ! MBPT tables as loaded: d = (Ksig, Kdsig, NmaxS, LmaxS, NhintS)
    Subroutine pconf_sigma_dims(d) bind(C, name="pconf_sigma_dims")
        Integer(c_int32_t), Intent(Out) :: d(5)
        d = [Ksig, Kdsig, NmaxS, LmaxS, NhintS]
    End Subroutine pconf_sigma_dims

    Subroutine pconf_sigma_tables(iint1s_o, rsig_o, dsig_o, esig_o, eps_o, diag_o) bind(C, name="pconf_sigma_tables")
        Integer(c_int32_t), Intent(Out) :: iint1s_o(*)
        Real(c_double), Intent(Out) :: rsig_o(*), dsig_o(*), esig_o(*), eps_o(*), diag_o(*)
        If (NhintS > 0) Then
            iint1s_o(1:NhintS) = Iint1S; rsig_o(1:NhintS) = Rsig; dsig_o(1:NhintS) = Dsig; esig_o(1:NhintS) = Esig
        End If
        eps_o(1:Ns) = Eps(1:Ns)
        If (Allocated(DiagS)) diag_o(1:Nd) = DiagS
    End Subroutine pconf_sigma_tables

    ! The valence energy of a Kdsig = 1 run, given before pconf_load instead
    ! of on standard input
    ! configuration classes of a kCSF > 0 space (nc_neq_o(Nc)); returns Nconf_neq
    Integer(c_int) Function pconf_csf_classes(nc_neq_o) bind(C, name="pconf_csf_classes")
        Integer(c_int), Intent(Out) :: nc_neq_o(*)
        pconf_csf_classes = Nconf_neq
        If (Allocated(Nc_neq) .and. Nconf_neq > 0) nc_neq_o(1:Nc) = Nc_neq
    End Function pconf_csf_classes

    Subroutine pconf_set_e0(e0) bind(C, name="pconf_set_e0")
        Real(c_double), Value :: e0
        E_0 = e0
        have_e0 = .true.
    End Subroutine pconf_set_e0

    Subroutine pconf_error(buf, n) bind(C, name="pconf_error")
        Character(Kind=c_char), Intent(Out) :: buf(*)
        Integer(c_int), Value :: n
        Integer :: i, l
        l = Min(Len_Trim(host_err), n - 1)
        Do i = 1, l
            buf(i) = host_err(i:i)
        End Do
        buf(l + 1) = c_null_char
    End Subroutine pconf_error

End Module pconf_host
