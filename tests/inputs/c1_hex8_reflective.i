# C1: 8x8x8 HEX8 reflective 10 cm cube, one A=12 nuclide (0.08 /b-cm),
# isotropic 2 MeV point source at the centre, per-element track-length and
# collision flux tallies.  Laid out as test/tests/study/openmc_study.i.
[Mesh]
  [gmg]
    type = GeneratedMeshGenerator
    dim = 3
    nx = 8
    ny = 8
    nz = 8
    xmin = -5
    ymin = -5
    zmin = -5
    xmax = 5
    ymax = 5
    zmax = 5
  []
[]

[Problem]
  solve = false
  kernel_coverage_check = false
[]

[Variables/temperature]
  order = CONSTANT
  family = MONOMIAL
[]

[AuxVariables]
  [flux_tl]
    order = CONSTANT
    family = MONOMIAL
  []
  [flux_coll]
    order = CONSTANT
    family = MONOMIAL
  []
[]

[RayKernels/collision]
  type = CollisionKernel
  temperature = temperature
  blocks = "0"
  materials = "1"
[]

[RayBCs]
  [reflect]
    type = ReflectRayBC
    boundary = 'back front top right left bottom'
  []
[]

[MonteCarlo]
  particles = 10000
  batches = 10
  seed = 1
  [source]
    space = point
    point = '0 0 0'
    energy = 2e6
  []
  [Nuclides]
    [C12]
      type = SyntheticNuclide
      awr = 12
      n_points = 20000
      seed = 202
      sigma_a0 = 1.0
      sigma_p = 4.7
      resonances = 50
    []
  []
  [Materials]
    [graphite]
      id = 1
      nuclides = 'C12'
      densities = '0.08'
    []
  []
[]

[UserObjects]
  [study]
    type = OpenMCStudy
    execute_on = TIMESTEP_END
    verbose = false
  []
  [tracklength]
    type = OpenMCTally
    id = 1
    particle_type = 'neutron'
    estimator = 'TRACKLENGTH'
    scores = 'flux'
    filters = 'cell'
    execute_on = 'initial'
  []
  [collision]
    type = OpenMCTally
    id = 2
    particle_type = 'neutron'
    estimator = 'COLLISION'
    scores = 'flux'
    filters = 'cell'
    execute_on = 'initial'
  []
[]

[AuxKernels]
  [tl]
    type = OpenMCTallyAux
    tally_id = 1
    granularity = cell
    estimator = TRACKLENGTH
    particle_type = neutron
    score = 'flux'
    variable = flux_tl
  []
  [coll]
    type = OpenMCTallyAux
    tally_id = 2
    granularity = cell
    estimator = COLLISION
    particle_type = neutron
    score = 'flux'
    variable = flux_coll
  []
[]

[Executioner]
  type = Transient
[]

[Outputs]
  csv = true
[]
