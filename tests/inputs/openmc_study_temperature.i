# The eigenvalue study of openmc_study_eigen.i with element temperatures: the
# collision kernel's CONSTANT MONOMIAL temperature variable is 600 K (its
# initial condition) in the fuel block and 1150 K (a ConstantIC) in the centre
# block, and U235 and O16 are tabulated at four temperatures; OpenMC reads the
# table nearest in kT per element (600 K and 1200 K here).
[Mesh]
  [gmg]
    type = GeneratedMeshGenerator
    dim = 3
    nx = 5
    ny = 5
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
    top_right = '1 1 1'
    bottom_left = '-1 -1 -1'
    block_id = 1
    block_name = 'center'
  []
[]

[Problem]
  solve = false
  kernel_coverage_check = false
[]

[Variables/temperature]
  order = CONSTANT
  family = MONOMIAL
  initial_condition = 600
[]

[ICs/centre_temperature]
  type = ConstantIC
  variable = temperature
  value = 1150
  block = 1
[]

[AuxVariables/power]
  order = CONSTANT
  family = MONOMIAL
[]

[RayKernels/collision]
  type = CollisionKernel
  temperature = temperature
  blocks = "0 1"
  materials = "1 2"
[]

[RayBCs]
  [reflect]
    type = ReflectRayBC
    boundary = 'back front top right left bottom'
  []
[]

[MonteCarlo]
  run_mode = eigenvalue
  particles = 30
  batches = 3
  inactive = 1
  seed = 1
  [source]
    space = box
    lower_left = '-0.62992 -0.62992 -0.62992'
    upper_right = '0.62992 0.62992 0.62992'
    watt_a = 0.988e6
    watt_b = 2.249e-6
  []
  [Nuclides]
    [U235]
      type = SyntheticNuclide
      awr = 233.02
      n_points = 20000
      seed = 235
      sigma_a0 = 680
      sigma_a_flat = 1.5
      sigma_p = 10
      fission_fraction = 0.85
      nu = '2.43'
      temperatures = '293.6 600 900 1200'
    []
    [O16]
      type = SyntheticNuclide
      awr = 15.86
      n_points = 20000
      seed = 16
      sigma_a0 = 0.0002
      sigma_p = 3.8
      temperatures = '293.6 600 900 1200'
    []
    [H1]
      type = SyntheticNuclide
      awr = 0.999
      n_points = 20000
      seed = 1
      sigma_a0 = 0.33
      sigma_p = 20
      resonances = 0
    []
  []
  [Materials]
    [fuel]
      id = 1
      nuclides = 'U235 O16 H1'
      densities = '0.02 0.04 0.06'
    []
    [center]
      id = 2
      nuclides = 'O16 H1'
      densities = '0.03 0.06'
    []
  []
[]

[UserObjects]
  [study]
    type = OpenMCStudy
    execute_on = TIMESTEP_END
  []
  [power]
    type = OpenMCTally
    id = 1
    particle_type = 'neutron'
    estimator = 'TRACKLENGTH'
    scores = 'fission'
    filters = 'cell'
    execute_on = 'initial'
  []
[]

[AuxKernels/power]
  type = OpenMCTallyAux
  tally_id = 1
  granularity = cell
  estimator = TRACKLENGTH
  particle_type = neutron
  score = 'fission'
  variable = power
[]

[Executioner]
  type = Transient
[]

[Outputs]
  csv = true
[]
