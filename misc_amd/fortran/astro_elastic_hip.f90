! astro_elastic_hip.f90 -- GPU element assembly for a Fortran host.
!
! gpuGetElasticSolnLoad(buildMat) has the effect of the reference's
! getElasticSolnLoad (Working/AStrO_commandFunctions.f90:260-351) on the
! reference's own arrays (module AStrO_globalData): with buildMat = 1 the
! element tangents are assembled into elasticMat with elMatCols, and in any
! case elasticLoad(k) is decreased by the internal force R(u) of the current
! nodeDisp (R = K u, the linear static tangent at rest).  The work runs in
! misc_amd/libastro_hip.so (include/astro_abi.h) through ISO_C_BINDING: the
! first call builds a device context from the host's model arrays --
! nodeList, elementList, elementType, the sections, layups and materials
! with the design-variable adjustments of r_getMaterialStiffness,
! r_getMaterialDensity and r_getElThickOffset / r_getShellStiffness
! (Working/AStrO_r_designPropertyFunctions.f90:8-103, 223-330), and the
! constraint terms with their node or node-set names -- and later calls reuse
! it until the design vector r_dVec changes.
!
! gpuGetdRelasticdD(time, dVarNum) has the effect of the reference's
! getdRelasticdD (Working/AStrO_commandFunctions.f90:1172-1375) for brick8
! meshes: dRudD(1:elMatDim) = dR/dd of variable dVarNum at the current
! nodeDisp, with R = K u - F.  The host's design maps (dCategory, dSubCat,
! dComponent, elToDRange / elToD / elToCoef, ndToDRange / ndToD / ndToCoef)
! go to astro_ctx_set_design_vars; all variables' columns are formed on the
! device by complex step in one call (astro_ctx_dRdD_vars) and kept until
! nodeDisp or time changes, so the host's per-variable loop (getTotaldLdD)
! reads one column per call.  Element terms scatter into the element's own
! DOFs (the reference's D9 indexing is not reproduced).  Categories: elastic
! modulus / poissonRatio / shearModulus, density, thermalCond, nodeCoord and
! elastic nodeLoad (active when time lies in its activeTime window); another
! category stops the run when its column is asked for.
!
! Scope: brick8 (8) and shell4 (41) meshes of one element type; isotropic or
! orthotropic engineering constants (a material given by its 21-entry
! stiffness matrix stops the run).  For shell4 the condensed element matrix
! and the nodal load are produced; the internal-mode arrays intElasticMat and
! intElasticLoad are left as they are (zero for a linear step from rest).
! A device or argument error stops the run with the library's message.
module AStrO_gpuElastic
    use, intrinsic :: iso_c_binding
    use AStrO_constantVals
    use AStrO_globalData
    implicit none
    private
    public :: gpuGetElasticSolnLoad, gpuGetdRelasticdD, gpuElasticRelease

    type, bind(C) :: astro_model_t
        integer(c_int32_t) :: numNodes, numEls
        type(c_ptr) :: nodeList, elementList, elementType, eprops, density, cond
        integer(c_int32_t) :: numMpcTerms
        type(c_ptr) :: mpcEqn, mpcDof, mpcCoef, mpcIsSet, mpcNodeStart, mpcNodes
        integer(c_int32_t) :: numMPC
        type(c_ptr) :: mpcRHS
        integer(c_int32_t) :: numLoads
        type(c_ptr) :: loadNode, loadVal
        integer(c_int32_t) :: numLayers
        type(c_ptr) :: layers, zOffset, orient
        type(c_ptr) :: specHeat
        type(c_ptr) :: thermExp
    end type

    type, bind(C) :: astro_design_t
        integer(c_int32_t) :: numDV
        type(c_ptr) :: category, component, elToDRange, elToD, elToCoef, ndToDRange, ndToD, ndToCoef
        real(c_double) :: h
    end type

    type, bind(C) :: astro_options_t
        integer(c_int32_t) :: maxBW, literalMaxBW, thermLiteralD8, numVecs
        real(c_double) :: rtol
        integer(c_int32_t) :: maxCycles, coupledProfile, precision, stepGraphs
    end type

    interface
        integer(c_int) function astro_ctx_create(model, opts, ctx) bind(C, name='astro_ctx_create')
            import :: c_int, c_ptr, astro_model_t, astro_options_t
            type(astro_model_t), intent(in) :: model
            type(astro_options_t), intent(in) :: opts
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function astro_ctx_destroy(ctx) bind(C, name='astro_ctx_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function astro_ctx_assemble_elastic(ctx, massFactor) bind(C, name='astro_ctx_assemble_elastic')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: massFactor
        end function
        integer(c_int) function astro_ctx_elastic_soln_load(ctx) bind(C, name='astro_ctx_elastic_soln_load')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function astro_ctx_get_f64(ctx, name, host, count) bind(C, name='astro_ctx_get_f64')
            import :: c_int, c_ptr, c_char, c_double, c_int64_t
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: name(*)
            real(c_double), intent(out) :: host(*)
            integer(c_int64_t), value :: count
        end function
        integer(c_int) function astro_ctx_get_i32(ctx, name, host, count) bind(C, name='astro_ctx_get_i32')
            import :: c_int, c_ptr, c_char, c_int32_t, c_int64_t
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: name(*)
            integer(c_int32_t), intent(out) :: host(*)
            integer(c_int64_t), value :: count
        end function
        integer(c_int) function astro_ctx_set_f64(ctx, name, host, count) bind(C, name='astro_ctx_set_f64')
            import :: c_int, c_ptr, c_char, c_double, c_int64_t
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: name(*)
            real(c_double), intent(in) :: host(*)
            integer(c_int64_t), value :: count
        end function
        integer(c_int) function astro_ctx_set_design_vars(ctx, design) bind(C, name='astro_ctx_set_design_vars')
            import :: c_int, c_ptr, astro_design_t
            type(c_ptr), value :: ctx
            type(astro_design_t), intent(in) :: design
        end function
        integer(c_int) function astro_ctx_dRdD_vars(ctx, dRdD) bind(C, name='astro_ctx_dRdD_vars')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: dRdD(*)
        end function
        type(c_ptr) function astro_last_error() bind(C, name='astro_last_error')
            import :: c_ptr
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

    type(c_ptr), save :: ctx = c_null_ptr
    logical, save :: assembled = .false.
    real*8, allocatable, save :: dVecAtSetup(:)
    ! design maps handed to the context, and the dR/dd columns of the last
    ! (nodeDisp, time) they were formed for
    logical, save :: designSet = .false., haveCols = .false.
    integer, allocatable, save :: dvKind(:)
    real*8, allocatable, save :: dRdDCols(:), uAtCols(:)
    real*8, save :: tAtCols = 0d0

contains

    subroutine check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: n, i
        character(len=512) :: text
        if (rc .eq. 0) return
        n = int(c_strlen(astro_last_error()))
        call c_f_pointer(astro_last_error(), msg, [n])
        text = ' '
        do i = 1, min(n, 512)
            text(i:i) = msg(i)
        enddo
        write(*,'(A,A,A,I0,A,A)') 'gpuElastic: ', what, ' failed (', rc, '): ', trim(text)
        stop 1
    end subroutine check

    ! Node list of a constraint term: a node number, or the nodes of a node set.
    subroutine term_nodes(name, nodes, cnt, isSet)
        character(len=*), intent(in) :: name
        integer(c_int32_t), intent(out) :: nodes(:)
        integer, intent(out) :: cnt, isSet
        integer :: nd, ios, q, k
        read(name, *, iostat=ios) nd
        if (ios .eq. 0) then
            nodes(1) = nd
            cnt = 1
            isSet = 0
            return
        endif
        isSet = 1
        cnt = 0
        do q = 1, numNdSets
            if (ndSetName(q) .eq. name) then
                do k = ndSetRange(q-1) + 1, ndSetRange(q)
                    cnt = cnt + 1
                    nodes(cnt) = nodeSets(k)
                enddo
                return
            endif
        enddo
        write(*,'(A,A)') 'gpuElastic: unknown node or node set in a constraint: ', trim(name)
        stop 1
    end subroutine term_nodes

    ! The model arrays in the library's layout and the device context.
    subroutine setup()
        type(astro_model_t) :: m
        type(astro_options_t) :: o
        real(c_double), allocatable, target :: xyz(:), ep(:), den(:), cnd(:), lay(:), zo(:), ori(:), coef(:), rhs(:)
        real(c_double), allocatable, target :: sh(:), tex(:)
        integer(c_int32_t), allocatable, target :: conn(:), et(:), eqn(:), dof(:), isSet(:), nstart(:), nodes(:)
        integer(c_int32_t), allocatable :: tn(:)
        integer :: e, sec, mat, i, j, l, d, nL, nLay, cnt, iset, t, nTerms, k
        real*8 :: lp(6)

        if (c_associated(ctx)) then
            if (numDVar .eq. 0) return
            if (all(r_dVec .eq. dVecAtSetup)) return
            call gpuElasticRelease()
        endif
        allocate(tn(max(1, numNodes)))
        allocate(xyz(3*numNodes), conn(8*numEls), et(numEls), ep(9*numEls), den(numEls), cnd(6*numEls), sh(numEls), &
                 tex(6*numEls))
        xyz = reshape(nodeList(1:3, 1:numNodes), [3*numNodes])
        ! node coordinates with their design variables (r_getNodeCoord)
        if (numDVar .gt. 0) then
            do i = 1, numNodes
                do j = ndToDRange(i-1) + 1, ndToDRange(i)
                    d = ndToD(j)
                    if (d .gt. 0) then
                        if (dCategory(d) .eq. 'nodeCoord') xyz(3*(i-1)+dComponent(d)) = &
                            xyz(3*(i-1)+dComponent(d)) + ndToCoef(j)*r_dVec(d)
                    endif
                enddo
            enddo
        endif
        conn = reshape(elementList(1:8, 1:numEls), [8*numEls])
        et = elementType(1:numEls)
        nLay = 0
        do e = 1, numEls
            sec = elementSection(e)
            if (elementType(e) .eq. 41) nLay = max(nLay, secLayupRange(sec) - secLayupRange(sec-1))
        enddo
        allocate(lay(max(1, 6*nLay*numEls)), zo(numEls), ori(6*numEls))
        lay = r_0
        zo = r_0
        ori = r_0
        do e = 1, numEls
            sec = elementSection(e)
            mat = sectionMatId(sec)
            if (elementType(e) .eq. 41) mat = layupMatId(secLayupRange(sec-1) + 1)
            if (abs(materialStiffMat(1, mat)) .gt. 0d0) then
                write(*,'(A)') 'gpuElastic: materials given by a stiffness matrix are not supported'
                stop 1
            endif
            ! engineering constants and density with the element's design
            ! variables (r_getMaterialStiffness, r_getMaterialDensity)
            ep(9*(e-1)+1:9*e) = materialElastic(1:9, mat)
            den(e) = materialDensity(mat)
            cnd(6*(e-1)+1:6*e) = materialThermCond(1:6, mat)
            sh(e) = materialSpecHeat(mat)
            tex(6*e-5:6*e) = materialThermExp(1:6, mat)
            if (numDVar .gt. 0) then
                do i = elToDRange(e-1) + 1, elToDRange(e)
                    d = elToD(i)
                    if (dCategory(d) .eq. 'elastic') then
                        if (dSubCat(d) .eq. 'modulus') then
                            j = dComponent(d)
                        elseif (dSubCat(d) .eq. 'poissonRatio') then
                            j = dComponent(d) + 3
                        else
                            j = dComponent(d) + 6
                        endif
                        ep(9*(e-1)+j) = ep(9*(e-1)+j) + elToCoef(i)*r_dVec(d)
                    elseif (dCategory(d) .eq. 'density') then
                        den(e) = den(e) + elToCoef(i)*r_dVec(d)
                    endif
                enddo
            endif
            if (elementType(e) .eq. 41) then
                ! layers (E1, E2, nu12, G12, thickness, angle) and the z offset
                ! (r_getShellStiffness, r_getElThickOffset)
                nL = 0
                do l = secLayupRange(sec-1) + 1, secLayupRange(sec)
                    nL = nL + 1
                    j = layupMatId(l)
                    lp = (/ materialElastic(1, j), materialElastic(2, j), materialElastic(4, j), &
                            materialElastic(7, j), layupThickness(l), layupAngle(l) /)
                    if (numDVar .gt. 0) then
                        do i = elToDRange(e-1) + 1, elToDRange(e)
                            d = elToD(i)
                            if (dCategory(d) .eq. 'layup' .and. dLayer(d) .eq. nL) then
                                if (dSubCat(d) .eq. 'modulus') then
                                    k = dComponent(d)
                                    lp(k) = lp(k) + elToCoef(i)*r_dVec(d)
                                elseif (dSubCat(d) .eq. 'poissonRatio') then
                                    lp(3) = lp(3) + elToCoef(i)*r_dVec(d)
                                elseif (dSubCat(d) .eq. 'shearModulus') then
                                    lp(4) = lp(4) + elToCoef(i)*r_dVec(d)
                                elseif (dSubCat(d) .eq. 'thickness') then
                                    lp(5) = lp(5) + elToCoef(i)*r_dVec(d)
                                elseif (dSubCat(d) .eq. 'angle') then
                                    lp(6) = lp(6) + elToCoef(i)*r_dVec(d)
                                endif
                            endif
                        enddo
                    endif
                    lay(6*nLay*(e-1) + 6*(nL-1) + 1 : 6*nLay*(e-1) + 6*nL) = lp
                enddo
                zo(e) = sectionZOffset(sec)
                if (numDVar .gt. 0) then
                    do i = elToDRange(e-1) + 1, elToDRange(e)
                        d = elToD(i)
                        if (dCategory(d) .eq. 'layup' .and. dSubCat(d) .eq. 'zOffset') &
                            zo(e) = zo(e) + elToCoef(i)*r_dVec(d)
                    enddo
                endif
                ori(6*(e-1)+1:6*(e-1)+3) = sectionOrient(1:3, 3*sec-2)
                ori(6*(e-1)+4:6*e) = sectionOrient(1:3, 3*sec-1)
            endif
        enddo
        ! constraint terms with their node lists
        nTerms = 0
        if (allocated(mpcEqn)) nTerms = size(mpcEqn)
        allocate(eqn(max(1, nTerms)), dof(max(1, nTerms)), coef(max(1, nTerms)), isSet(max(1, nTerms)), &
                 nstart(nTerms+1), rhs(max(1, numMPC)))
        cnt = 0
        do t = 1, nTerms
            call term_nodes(mpcNode(t), tn, i, iset)
            cnt = cnt + i
        enddo
        allocate(nodes(max(1, cnt)))
        nstart(1) = 0
        do t = 1, nTerms
            call term_nodes(mpcNode(t), tn, i, iset)
            eqn(t) = mpcEqn(t)
            dof(t) = mpcDof(t)
            coef(t) = mpcCoef(t)
            isSet(t) = iset
            nodes(nstart(t)+1:nstart(t)+i) = tn(1:i)
            nstart(t+1) = nstart(t) + i
        enddo
        if (numMPC .gt. 0) rhs(1:numMPC) = mpcRHS(1:numMPC)

        m%numNodes = numNodes
        m%numEls = numEls
        m%nodeList = c_loc(xyz)
        m%elementList = c_loc(conn)
        m%elementType = c_loc(et)
        m%eprops = c_loc(ep)
        m%density = c_loc(den)
        m%cond = c_loc(cnd)
        m%numMpcTerms = nTerms
        m%mpcEqn = c_loc(eqn)
        m%mpcDof = c_loc(dof)
        m%mpcCoef = c_loc(coef)
        m%mpcIsSet = c_loc(isSet)
        m%mpcNodeStart = c_loc(nstart)
        m%mpcNodes = c_loc(nodes)
        m%numMPC = numMPC
        m%mpcRHS = c_loc(rhs)
        m%numLoads = 0
        m%loadNode = c_null_ptr
        m%loadVal = c_null_ptr
        m%specHeat = c_loc(sh)
        m%thermExp = c_loc(tex)
        m%numLayers = nLay
        if (nLay .gt. 0) then
            m%layers = c_loc(lay)
            m%zOffset = c_loc(zo)
            m%orient = c_loc(ori)
        else
            m%layers = c_null_ptr
            m%zOffset = c_null_ptr
            m%orient = c_null_ptr
        endif
        ! the host's own solverMaxBW and its 32-bit rule (the profile is not
        ! used here; the CSR structure is the same either way)
        o%maxBW = solverMaxBW
        o%literalMaxBW = 1
        o%thermLiteralD8 = 1
        o%numVecs = 0
        o%rtol = 0d0
        o%maxCycles = 0
        o%coupledProfile = 0
        o%precision = 0
        o%stepGraphs = 0
        call check(astro_ctx_create(m, o, ctx), 'astro_ctx_create')
        assembled = .false.
        designSet = .false.
        haveCols = .false.
        if (numDVar .gt. 0) then
            if (allocated(dVecAtSetup)) deallocate(dVecAtSetup)
            allocate(dVecAtSetup(numDVar))
            dVecAtSetup = r_dVec
        endif
    end subroutine setup

    subroutine gpuGetElasticSolnLoad(buildMat)
        integer, intent(in) :: buildMat
        real(c_double), allocatable :: ld(:)
        call setup()
        if (buildMat .eq. 1 .or. .not. assembled) then
            call check(astro_ctx_assemble_elastic(ctx, 0d0), 'astro_ctx_assemble_elastic')
            assembled = .true.
        endif
        if (buildMat .eq. 1) then
            call check(astro_ctx_get_f64(ctx, 'elasticMat'//c_null_char, elasticMat, int(elMatSize, c_int64_t)), &
                       'reading elasticMat')
            call check(astro_ctx_get_i32(ctx, 'elMatCols'//c_null_char, elMatCols, int(elMatSize, c_int64_t)), &
                       'reading elMatCols')
        endif
        allocate(ld(elMatDim))
        call check(astro_ctx_set_f64(ctx, 'nodeDisp'//c_null_char, nodeDisp, int(elMatDim, c_int64_t)), &
                   'writing nodeDisp')
        ! the element stress holds C (strain - T alpha): the host's nodal
        ! temperatures (rank-indexed, as nodeTemp on the device)
        if (any(materialThermExp .ne. 0d0)) &
            call check(astro_ctx_set_f64(ctx, 'nodeTemp'//c_null_char, nodeTemp, int(numNodes, c_int64_t)), &
                       'writing nodeTemp')
        call check(astro_ctx_elastic_soln_load(ctx), 'astro_ctx_elastic_soln_load')
        call check(astro_ctx_get_f64(ctx, 'elasticLoad'//c_null_char, ld, int(elMatDim, c_int64_t)), &
                   'reading elasticLoad')
        elasticLoad(1:elMatDim) = elasticLoad(1:elMatDim) + ld
    end subroutine gpuGetElasticSolnLoad

    ! The host's design maps in the library's form: category codes of
    ! include/astro_abi.h (0 where the library has no counterpart; such a
    ! variable enters as a density variable without map entries, which moves
    ! nothing in the static residual), the node map with the coefficients of
    ! elastic nodeLoad variables outside their activeTime window set to 0.
    subroutine design_setup(time)
        real*8, intent(in) :: time
        type(astro_design_t) :: dsn
        integer(c_int32_t), allocatable, target :: cat(:), comp(:), er(:), eid(:), nr(:), nid(:)
        real(c_double), allocatable, target :: ec(:), nc(:)
        integer :: k, j, ne, nn
        if (allocated(dvKind)) deallocate(dvKind)
        allocate(dvKind(numDVar), cat(numDVar), comp(numDVar))
        do k = 1, numDVar
            dvKind(k) = 0
            if (dCategory(k) .eq. 'elastic') then
                if (dSubCat(k) .eq. 'modulus') dvKind(k) = 1
                if (dSubCat(k) .eq. 'poissonRatio') dvKind(k) = 2
                if (dSubCat(k) .eq. 'shearModulus') dvKind(k) = 3
            elseif (dCategory(k) .eq. 'density') then
                dvKind(k) = 4
            elseif (dCategory(k) .eq. 'thermalCond') then
                dvKind(k) = 5
            elseif (dCategory(k) .eq. 'nodeCoord') then
                dvKind(k) = 6
            elseif (dCategory(k) .eq. 'nodeLoad' .and. dSubCat(k) .eq. 'elastic') then
                dvKind(k) = 7
            endif
            cat(k) = dvKind(k)
            comp(k) = dComponent(k)
            if (dvKind(k) .eq. 0) then
                cat(k) = 4
                comp(k) = 1
            endif
        enddo
        ne = elToDRange(numEls)
        nn = ndToDRange(numNodes)
        allocate(er(numEls+1), eid(max(1, ne)), ec(max(1, ne)), nr(numNodes+1), nid(max(1, nn)), nc(max(1, nn)))
        er(1) = 0
        j = 0
        do k = 1, numEls
            do ne = elToDRange(k-1) + 1, elToDRange(k)
                if (elToD(ne) .gt. 0) then
                    if (dvKind(elToD(ne)) .ne. 0) then
                        j = j + 1
                        eid(j) = elToD(ne)
                        ec(j) = elToCoef(ne)
                    endif
                endif
            enddo
            er(k+1) = j
        enddo
        nr(1) = 0
        j = 0
        do k = 1, numNodes
            do nn = ndToDRange(k-1) + 1, ndToDRange(k)
                if (ndToD(nn) .gt. 0) then
                    if (dvKind(ndToD(nn)) .ne. 0) then
                        j = j + 1
                        nid(j) = ndToD(nn)
                        nc(j) = ndToCoef(nn)
                        if (dvKind(ndToD(nn)) .eq. 7) then
                            if (time .lt. dActTime(1, ndToD(nn)) .or. time .gt. dActTime(2, ndToD(nn))) nc(j) = 0d0
                        endif
                    endif
                endif
            enddo
            nr(k+1) = j
        enddo
        dsn%numDV = numDVar
        dsn%category = c_loc(cat)
        dsn%component = c_loc(comp)
        dsn%elToDRange = c_loc(er)
        dsn%elToD = c_loc(eid)
        dsn%elToCoef = c_loc(ec)
        dsn%ndToDRange = c_loc(nr)
        dsn%ndToD = c_loc(nid)
        dsn%ndToCoef = c_loc(nc)
        dsn%h = 1d-30
        call check(astro_ctx_set_design_vars(ctx, dsn), 'astro_ctx_set_design_vars')
        designSet = .true.
        haveCols = .false.
    end subroutine design_setup

    subroutine gpuGetdRelasticdD(time, dVarNum)
        real*8, intent(in) :: time
        integer, intent(in) :: dVarNum
        integer :: n
        call setup()
        n = elMatDim
        if (.not. designSet .or. time .ne. tAtCols) call design_setup(time)
        if (dvKind(dVarNum) .eq. 0) then
            write(*,'(A,A,A,A)') 'gpuElastic: getdRelasticdD has no device form for design category ', &
                trim(dCategory(dVarNum)), ' / ', trim(dSubCat(dVarNum))
            stop 1
        endif
        if (haveCols) then
            if (any(nodeDisp(1:n) .ne. uAtCols)) haveCols = .false.
        endif
        if (.not. haveCols) then
            if (allocated(dRdDCols)) deallocate(dRdDCols)
            if (allocated(uAtCols)) deallocate(uAtCols)
            allocate(dRdDCols(n*numDVar), uAtCols(n))
            uAtCols = nodeDisp(1:n)
            call check(astro_ctx_set_f64(ctx, 'nodeDisp'//c_null_char, nodeDisp, int(n, c_int64_t)), &
                       'writing nodeDisp')
            call check(astro_ctx_dRdD_vars(ctx, dRdDCols), 'astro_ctx_dRdD_vars')
            tAtCols = time
            haveCols = .true.
        endif
        dRudD(1:n) = dRdDCols((dVarNum-1)*n+1 : dVarNum*n)
    end subroutine gpuGetdRelasticdD

    ! Frees the device context (a later call builds it again).
    subroutine gpuElasticRelease()
        integer(c_int) :: rc
        if (c_associated(ctx)) rc = astro_ctx_destroy(ctx)
        ctx = c_null_ptr
        assembled = .false.
        designSet = .false.
        haveCols = .false.
    end subroutine gpuElasticRelease

end module AStrO_gpuElastic
