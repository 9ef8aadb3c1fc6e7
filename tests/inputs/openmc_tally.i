# Tally user objects and their read-out, laid out as
# test/tests/tally/openmc_tally.i: a 10x10x1 box with a centre block,
# universe and cell tallies with several scores and an energy filter.
[Mesh]
  [gmg]
    type = GeneratedMeshGenerator
    dim = 3
    nx = 10
    ny = 10
    nz = 1
    xmin = -5
    ymin = -5
    zmin = -5
    xmax = 5
    ymax = 5
    zmax = 5
  []
  [add_subdomain]
    input = gmg
    type = SubdomainBoundingBoxGenerator
    top_right = '3 3 5'
    bottom_left = '-3 -3 -5'
    block_id = 1
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
  [tally]
    order = CONSTANT
    family = MONOMIAL
  []
  [cell_flux]
    order = CONSTANT
    family = MONOMIAL
  []
[]

[RayKernels/collision]
  type = CollisionKernel
  temperature = temperature
  # mesh block ids
  blocks = "0 1"
  # material ids
  materials = "4 1"
[]

[RayBCs]
  [reflect]
    type = ReflectRayBC
    boundary = 'back front top right left bottom'
  []
[]

[MonteCarlo]
  particles = 2000
  batches = 3
  [source]
    space = point
    point = '0.1 0.2 0.3'
    energy = 2e6
  []
  [Nuclides]
    [H1]
      awr = 1
      n_points = 3000
      seed = 11
      sigma_a0 = 0.33
      sigma_p = 20
    []
    [O16]
      awr = 16
      n_points = 3000
      seed = 12
      sigma_a0 = 0.0002
      sigma_p = 3.8
    []
    [Fe56]
      awr = 56
      n_points = 2000
      seed = 13
      sigma_a0 = 2.56
      sigma_p = 11
    []
  []
  [Materials]
    [water]
      id = 1
      nuclides = 'H1 O16'
      densities = '0.0668 0.0334'
    []
    [steel]
      id = 4
      nuclides = 'Fe56'
      densities = '0.085'
    []
  []
[]

[UserObjects]
  active = 'study univtally celltally'

  [study]
    type = OpenMCStudy
    execute_on = TIMESTEP_END
    verbose = false
  []

  [univtally]
    type = OpenMCTally
    id = 1
    particle_type = 'neutron'
    estimator = 'COLLISION'
    scores = 'flux absorption'
    filters = 'energy universe'
    energy_bins = '1e-5 1 1e3 2e7'
    execute_on = 'initial'
  []

  [celltally]
    type = OpenMCTally
    id = 2
    particle_type = 'neutron'
    estimator = 'TRACKLENGTH'
    scores = 'flux total'
    filters = 'cell'
    nuclides = 'H1 Fe56'
    execute_on = 'initial'
  []
[]

[Executioner]
  type = Transient
[]

[Outputs]
  csv = true
[]

[AuxKernels]
  [univ_val]
    type = OpenMCTallyAux
    tally_id = 1
    granularity = universe
    estimator = COLLISION
    particle_type = neutron
    score = 'absorption'
    variable = tally
  []
  [cell_val]
    type = OpenMCTallyAux
    tally_id = 2
    granularity = cell
    estimator = TRACKLENGTH
    particle_type = neutron
    nuclide = 'H1'
    score = 'total'
    variable = cell_flux
  []
[]
