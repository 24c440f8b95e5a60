! Test driver around the reference's own Fortran modules (TEST INFRASTRUCTURE).
!
! Built by oracle/Makefile (target `ref`) together with params.f90,
! conf_variables.F90, wigner.f90, str_fmt.f90, utils.f90, conf_init.f90,
! determinants.f90 and integrals.f90 compiled from ud-pci/pCI's src/ where they
! lie.  pconf.F90 itself needs MPI (mpi_f08), which this image does not have,
! so the few lines of it the checks need -- the CONF.DAT/CONF.GNT reads of
! Input/Init (src/pconf.F90:382-406, 443-530), Hmltn (1653-1692), the npes=1
! FormH loop (1226-1290) and, for MBPT runs (Kl = 2, Ksig = 1, 2), the Kl/Ksig
! rule (345-356), the Eps read (663-667), RintS (SGC.CON with the screening
! coefficients and SCRC.CON, 715-840) and FormD (852-880) -- are restated
! here around calls into the
! reference modules (ReadConfInp, Rint, Dinit, Det_Number, Det_List,
! Rspq_phase1/2, CompD, Gint, Hint, Gaunt, Build*LUT/Hash, FJ3).
! Every entry point is bind(C) and returns 0 on success.
Module ref_driver
    Use iso_c_binding
    Use conf_variables
    Use determinants, Only : Dinit, Det_Number, Det_List, Rspq_phase1, Rspq_phase2, CompD, Jterm
    Use integrals, Only : Rint, Gint, Hint, Gaunt, BuildGauntLUT, BuildHintLUT, BuildGintHash, BuildHintSLUT, &
                          BuildGintSHash
    Use conf_init, Only : ReadConfInp, ReadConfigurations, ReadCiIn
    Use wigner, Only : FJ3
    Implicit None
Contains

    ! Run pconf's start-up on the files of the current directory.
    Integer(c_int) Function ref_load() bind(C, name="ref_load")
        Integer :: i, ni, nj, nnj, llj, jlj, kkj, ic, n, i0, nmin, imax, j, nf, ii
        Real(dp) :: d, c1, c2
        Real(dp), Dimension(4*IP6) :: pq
        Real(dp), Dimension(IP6) :: p
        Integer :: lrec

        ref_load = 0
        open(unit=11, status='UNKNOWN', file='CONF.RES')
        Call ReadConfInp
        Call ReadCiIn
        ! MBPT effective integrals only for Kl = 2 (src/pconf.F90:345-356)
        If (Kl == 2) Then
            Kl = 0
        Else
            Ksig = 0
        End If
        Call ReadConfigurations
        ! -- Init, CONF.DAT part (src/pconf.F90:443-530)
        c1 = 0.01d0
        mj = Int(2*dabs(Jm)+0.01d0)
        Inquire(IOLENGTH=lrec) d
        Open(12, file='CONF.DAT', status='OLD', access='DIRECT', recl=IP6*lrec)
        Read(12, rec=1) p;  pq(1:IP6) = p
        Read(12, rec=2) p;  pq(IP6+1:2*IP6) = p
        Read(12, rec=5) p;  pq(2*IP6+1:3*IP6) = p
        Read(12, rec=6) p;  pq(3*IP6+1:4*IP6) = p
        If (Ksig > 0) Then  ! HF energies (src/pconf.F90:470, 663-667)
            ii = Int(pq(3)+c1)
            Do ni = Nso+1, Int(pq(2)+c1)
                Read(12, rec=2*ni+7) p
                Eps(ni) = -p(ii+1)
            End Do
        End If
        Close(12)
        Allocate(Nvc(Nc), Nc0(Nc), Nq(Nsp), Nip(Nsp))
        Ns = Int(pq(2)+c1)
        nf = 20
        Do ni = 1, Ns
            nf = nf+1; Nn(ni) = Int(pq(nf)+c1)
            nf = nf+1; Ll(ni) = Int(pq(nf)+c1)
            nf = nf+3; c2 = dsign(c1, pq(nf)); Kk(ni) = Int(pq(nf)+c2)
            nf = nf+1; c2 = dsign(c1, pq(nf)); Jj(ni) = Int(pq(nf)+c2)
        End Do
        Nsu = 0
        Do nj = 1, Nsp
            i = Int(sign(1.d0, Qnl(nj)))
            d = dabs(Qnl(nj))+1.d-14
            d = 10.0*d
            nnj = Int(d)
            d = 10.0d0*(d-nnj)
            llj = Int(d)
            jlj = 2*llj+i
            kkj = -i*((jlj+1)/2)
            d = 100.0d0*(d-llj)
            Nq(nj) = Int(d+0.1d0)
            Do ni = 1, Ns
                If (nnj == Nn(ni) .and. Kk(ni) == kkj) Exit
            End Do
            If (ni > Ns) Then
                ref_load = 2
                Return
            End If
            Nip(nj) = ni
            If (Nsu < ni) Nsu = ni
        End Do
        Nst = 0
        Do ni = 1, Nsu
            imax = 2*Jj(ni)+1
            Do j = 1, imax, 2
                Nst = Nst+1
            End Do
        End Do
        n = 0; ic = 0; i0 = 0; i = 0
        nmin = Nso+1
        Do ni = nmin, Nsp
            i = i+1
            n = n+Nq(ni)
            If (n >= Ne) Then
                ic = ic+1
                Nvc(ic) = i
                Nc0(ic) = Nso+i0
                i0 = i0+i
                n = 0
                i = 0
            End If
        End Do
        ! -- integrals, determinants
        Call Rint
        Call Dinit
        Call Det_Number
        Call Det_List(Iarr)
        ! -- CONF.GNT (src/pconf.F90:382-406)
        Open(unit=16, file='CONF.GNT', status='OLD', form='UNFORMATTED')
        Read(16) Nlx, Ngaunt
        Allocate(num_gaunts_per_partial_wave(Nlx+1), In(Ngaunt), Gnt(Ngaunt))
        Read(16) num_gaunts_per_partial_wave
        Read(16) (In(i), i=1, Ngaunt)
        Read(16) (Gnt(i), i=1, Ngaunt)
        Close(16)
        Call BuildGauntLUT
        Call BuildHintLUT
        Call BuildGintHash
        If (Ksig > 0) Call sigma_tables
        Allocate(iconf1(Ne), iconf2(Ne))
        Close(11)
    End Function ref_load

    ! RintS (src/pconf.F90:715-840; the valence energy E_0 of a Kdsig run
    ! comes from the file E0.in instead of the terminal, and the pause after
    ! the screening note is left out), then BuildHintSLUT, BuildGintSHash and
    ! FormD (src/pconf.F90:852-880).
    Subroutine sigma_tables
        Integer :: na, nb, la, lb, ja, jb, na1, nb1, i, khot, nsh, n, ie, ke, k, k1
        Integer :: ns1, nso1, nsh1, khot1, nsx1, nsx2, nmax1, lmax1, nav, kbox, nrd
        Character(Len=4) :: idummy
        Real :: x, y, z
        Real(dp) :: xe
        If (Kdsig /= 0) Then
            Open(unit=19, file='E0.in', status='OLD')
            Read(19, *) E_0
            Close(19)
            Kexn = 1
        End If
        nsh = Nso+1
        Open(unit=18, file='SGC.CON', status='OLD')
        Read(18, '(7X,I3,5X,I1,24X,I1)') NmaxS, LmaxS, khot
        NhintS = 0
        Do na = nsh, NmaxS
            la = Ll(na); ja = Jj(na)
            If (la > LmaxS) Cycle
            Do nb = na, NmaxS
                lb = Ll(nb); jb = Jj(nb)
                If (lb /= la .or. jb /= ja) Cycle
                NhintS = NhintS+1
            End Do
        End Do
        Allocate(Iint1S(NhintS), Rsig(NhintS), Dsig(NhintS), Esig(NhintS))
        i = 0
        Do na = nsh, NmaxS
            la = Ll(na); ja = Jj(na)
            If (la > LmaxS) Cycle
            Do nb = na, NmaxS
                lb = Ll(nb); jb = Jj(nb)
                If (lb /= la .or. jb /= ja) Cycle
                Read(18, *) idummy, na1, nb1, x, y, z
                i = i+1
                Iint1S(i) = Nx*(na-nsh)+(nb-nsh+1)
                Rsig(i) = x
                Dsig(i) = y
                Esig(i) = Real(z, kind=type2_real)
            End Do
        End Do
        If (Ksig == 2) Then
            Allocate(Scr(10))
            Read(18, '(5X,I2,5X,F8.5)')
            Do k = 1, 10
                Read(18, '(5X,I2,5X,F8.5)') k1, Scr(k)
            End Do
        End If
        Close(18)
        Call BuildHintSLUT
        If (Ksig == 2) Then
            Open(unit=13, file='SCRC.CON', status='OLD', form='UNFORMATTED')
            Read(13) ns1, nso1, nsh1, khot1, nsx1, nsx2, Ksym
            Read(13) nmax1, lmax1, Kmax, nav, kbox
            Read(13) NgintS, nrd
            Allocate(Rint2S(NgintS), Iint2S(NgintS), Iint3S(NgintS), Dint2S(NgintS), Eint2S(NgintS), IntOrdS(nrd))
            If (nsx2 == 0) Then
                Nsum = 4*nav
            Else
                Nsum = 4*nmax1
            End If
            Read(13) (Rint2S(i), i=1, NgintS)
            Read(13) (Iint2S(i), i=1, NgintS)
            Read(13) (Iint3S(i), i=1, NgintS)
            Read(13) (IntOrdS(i), i=1, nrd)
            Read(13) (Dint2S(i), i=1, NgintS)
            Read(13) (Eint2S(i), i=1, NgintS)
            Close(13)
            Call BuildGintSHash
        End If
        If (Ksig*Kdsig /= 0) Then
            Allocate(Diag(Nd))
            Do n = 1, Nd
                xe = 0.d0
                Do i = 1, Ne
                    ie = Iarr(i, n)
                    ke = Nh(ie)
                    xe = xe+Eps(ke)
                End Do
                xe = E_0-xe
                Diag(n) = Real(xe, kind=type_real)
            End Do
        End If
    End Subroutine sigma_tables

    ! The reference's Jterm (determinants.f90:185-348): Njd, Jt, Njt.  Its
    ! table goes to units 6 and 11; unit 11 is sent to /dev/null here.
    Integer(c_int) Function ref_jterms(jt_o, njt_o, cap) bind(C, name="ref_jterms")
        Integer(c_int), Value :: cap
        Integer(c_int), Intent(Out) :: jt_o(cap)
        Integer(c_int64_t), Intent(Out) :: njt_o(cap)
        Logical :: op
        Integer :: n
        Inquire(unit=11, opened=op)
        If (.not. op) Open(unit=11, file='/dev/null', status='old')
        Call Jterm
        n = Min(Njd, cap)
        jt_o(1:n) = Jt(1:n)
        njt_o(1:n) = Njt(1:n)
        ref_jterms = Njd
    End Function ref_jterms

    Subroutine ref_dims(nd_o, ne_o, nc_o, nst_o, nsu_o) bind(C, name="ref_dims")
        Integer(c_int), Intent(Out) :: nd_o, ne_o, nc_o, nst_o, nsu_o
        nd_o = Nd; ne_o = Ne; nc_o = Nc; nst_o = Nst; nsu_o = Nsu
    End Subroutine ref_dims

    Subroutine ref_space(iarr_o, ndc_o, jz_o, nh_o) bind(C, name="ref_space")
        Integer(c_int), Intent(Out) :: iarr_o(Ne, Nd), ndc_o(Nc), jz_o(Nst), nh_o(Nst)
        iarr_o = Iarr(1:Ne, 1:Nd)
        ndc_o = Ndc(1:Nc)
        jz_o = Jz(1:Nst)
        nh_o = Nh(1:Nst)
    End Subroutine ref_space

    Real(c_double) Function ref_gint(a, b, c, d) bind(C, name="ref_gint")
        Integer(c_int), Value :: a, b, c, d
        ref_gint = Gint(a, b, c, d)
    End Function ref_gint

    Real(c_double) Function ref_hint(a, b) bind(C, name="ref_hint")
        Integer(c_int), Value :: a, b
        ref_hint = Hint(a, b)
    End Function ref_hint

    Real(c_double) Function ref_gaunt(k, j1, m1, j2, m2) bind(C, name="ref_gaunt")
        Integer(c_int), Value :: k
        Real(c_double), Value :: j1, m1, j2, m2
        ref_gaunt = Gaunt(k, j1, m1, j2, m2)
    End Function ref_gaunt

    Subroutine ref_rspq(d1, d2, res) bind(C, name="ref_rspq")
        Integer(c_int), Intent(In) :: d1(Ne), d2(Ne)
        Integer(c_int), Intent(Out) :: res(6)
        Integer, Allocatable, Dimension(:) :: id1, id2
        Integer :: is, nf, ii(3), jj(3)
        Allocate(id1(Ne), id2(Ne))
        id1 = d1; id2 = d2
        Call Rspq_phase1(id1, id2, is, nf, ii, jj)
        If (nf <= 2) Call Rspq_phase2(id1, id2, is, nf, ii, jj)
        res = (/ nf, is, ii(2), ii(3), jj(2), jj(3) /)
    End Subroutine ref_rspq

    ! Hmltn (src/pconf.F90:1653-1692) for Gj = 0, around the reference Gint/Hint.
    Real(dp) Function hm(idet, is, nf, i2, i1, j2, j1)
        Integer, Allocatable, Dimension(:) :: idet
        Integer :: is, nf, i1, i2, j1, j2, iq, jq
        Real(dp) :: t
        t = 0.d0
        Select Case(nf)
        Case(2)
            t = t+Gint(i2, j2, i1, j1)*is
        Case(1)
            Do iq = 1, Ne
                i1 = idet(iq)
                If (i1 /= j1) t = t+Gint(j2, i1, j1, i1)*is
            End Do
            t = t+Hint(j2, j1)*is
        Case(0)
            Do iq = 1, Ne
                i1 = idet(iq)
                Do jq = iq+1, Ne
                    j1 = idet(jq)
                    t = t+Gint(i1, j1, i1, j1)*is
                End Do
                t = t+Hint(i1, i1)*is
            End Do
        End Select
        hm = t
    End Function hm

    ! npes=1 FormH loop (src/pconf.F90:1226-1290) over rows r0..r1.
    ! counts(r1-r0+1) receives per-row stored counts; col/val may be sized 1
    ! when cap == 0 (count-only call).
    Integer(c_int64_t) Function ref_formh(r0, r1, counts, cap, col, val) bind(C, name="ref_formh")
        Integer(c_int), Value :: r0, r1
        Integer(c_int64_t), Value :: cap
        Integer(c_int64_t), Intent(Out) :: counts(r1-r0+1)
        Integer(c_int), Intent(Out) :: col(*)
        Real(c_double), Intent(Out) :: val(*)
        Integer, Allocatable, Dimension(:) :: idet1, idet2
        Integer :: n, k, k1, kx, ic, icomp, diff, iSign, iIdx(3), jIdx(3)
        Integer(c_int64_t) :: pos, rc
        Real(dp) :: tt
        Allocate(idet1(Ne), idet2(Ne))
        pos = 0
        Do n = r0, r1
            idet1 = Iarr(1:Ne, n)
            iconf1(1:Ne) = Nh(idet1(1:Ne))
            rc = 0
            k = 0
            Do ic = 1, Nc
                kx = Ndc(ic)
                If (k+kx > n) kx = n-k
                If (kx /= 0) Then
                    idet2 = Iarr(1:Ne, k+1)
                    iconf2(1:Ne) = Nh(idet2(1:Ne))
                    Call CompD(iconf1, iconf2, icomp)
                    If (icomp > 2) Then
                        k = k+kx
                    Else
                        Do k1 = 1, kx
                            k = k+1
                            idet2 = Iarr(1:Ne, k)
                            Call Rspq_phase1(idet1, idet2, iSign, diff, iIdx, jIdx)
                            If (diff <= 2) Then
                                Call Rspq_phase2(idet1, idet2, iSign, diff, iIdx, jIdx)
                                If (Kdsig /= 0 .and. diff <= 2) E_k = Diag(k)   ! src/pconf.F90:1275
                                tt = hm(idet1, iSign, diff, jIdx(3), iIdx(3), jIdx(2), iIdx(2))
                                If (tt /= 0.d0) Then
                                    pos = pos+1
                                    rc = rc+1
                                    If (cap > 0) Then
                                        col(pos) = k
                                        val(pos) = tt
                                    End If
                                End If
                            End If
                        End Do
                    End If
                End If
            End Do
            counts(n-r0+1) = rc
        End Do
        ref_formh = pos
    End Function ref_formh

    ! Gaun (src/breit.f90:108-128) evaluated with the reference FJ3.
    Real(c_double) Function ref_gaun(k, j1, m1, j2, m2) bind(C, name="ref_gaun")
        Integer(c_int), Value :: k
        Real(c_double), Value :: j1, m1, j2, m2
        Integer :: is
        Real(dp) :: xk, q
        is = dabs(m1+0.5d0)+0.1d0
        If (is /= (is/2)*2) Then
            is = -1
        Else
            is = 1
        End If
        xk = k
        q = m1-m2
        ref_gaun = is*dsqrt((2*j1+1)*(2*j2+1))*FJ3(j1, j2, xk, -m1, m2, q)*FJ3(j1, j2, xk, 0.5d0, -0.5d0, 0.d0)
    End Function ref_gaun

End Module ref_driver
