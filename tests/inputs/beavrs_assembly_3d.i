# The reference's BEAVRS 3.1% assembly example (examples/beavrs_assembly/
# run_assembly_3d.i with its settings.xml): the 2D QUAD4 assembly mesh
# 31enr_16_in.e (a copy of the example's mesh file: 17x17 pin cells, blocks
# 10004 10006 10010 10014 11006, outer side sets 1-4) moved to z = -100 and
# extruded into 4 layers of HEX8 over 400 cm, reflective on all sides;
# eigenvalue, 200000 particles, 5 batches, no inactive batch, fission source
# on the plane z = 100 (OpenMC <space type="fission">).  Collision-estimator
# kappa-fission per element is the coupled power.
#
# The [MonteCarlo] block stands in for materials.xml and the HDF5 library:
# synthetic nuclides, and atom densities [1/(b cm)] from materials.xml's
# compositions with each isotope mapped to its element's stand-in (U234 and
# U236 to U238, Cr, Ni, Mn to Fe56, Sn to Zr90, B11 left out).  Only U235 and
# U238 carry resonances (13 for U238), which puts k near 1.35.
[Mesh]
  [assembly]
    type = FileMeshGenerator
    file = '31enr_16_in.e'
  []
  [translate]
    type = TransformGenerator
    input = assembly
    transform = TRANSLATE
    vector_value = '0 0 -100'
  []
  [extrude]
    type = FancyExtruderGenerator
    input = 'translate'
    num_layers = '4'
    heights = '400'
    direction = '0 0 1'
    bottom_boundary = '8'
    top_boundary = '9'
  []
[]

[Problem]
  solve = false
  kernel_coverage_check = false
  skip_nl_system_check = true
[]

[RayKernels/collision]
  type = CollisionKernel
  temperature = 300
  # mesh block ids
  blocks = "10000 10001 10004 10006 10008 10009 10010 10013 10014 11006"
  # openmc material id
  materials = "1 2 5 7 9 10 11 14 15 7"
  z_coord = 100
[]

[RayBCs]
  [lattice_reflect]
    type = ReflectRayBC
    boundary = '1 2 3 4 8 9'
    warn_non_planar = false
  []
[]

[UserObjects]
  [study]
    type = OpenMCStudy
    execute_on = TIMESTEP_END
    always_cache_traces = false
    segments_on_cache_traces = false
    data_on_cache_traces = false
    aux_data_on_cache_traces = false
    warn_subdomain_hmax = false
    warn_non_planar = false
    tolerate_failure = true
  []
[]

[MonteCarlo]
  run_mode = eigenvalue
  particles = 200000
  batches = 5
  inactive = 0
  seed = 1
  [source]
    space = fission
    lower_left = '-0.6 -20.7 100'
    upper_right = '20.7 0.6 100'
    watt_a = 0.988e6
    watt_b = 2.249e-6
  []
  [Nuclides]
    [U235]
      awr = 233.02
      n_points = 20000
      seed = 235
      sigma_a0 = 680
      sigma_a_flat = 1.5
      sigma_p = 10
      fission_fraction = 0.85
      nu = '2.43'
      kappa = 193.4e6
    []
    [U238]
      awr = 236.0
      n_points = 20000
      seed = 238
      sigma_a0 = 2.7
      sigma_p = 9.0
      resonances = 13
    []
    [O16]
      awr = 15.86
      n_points = 20000
      seed = 16
      sigma_a0 = 0.0002
      sigma_p = 3.8
      resonances = 0
    []
    [H1]
      awr = 0.999
      n_points = 20000
      seed = 1
      sigma_a0 = 0.33
      sigma_p = 20
      resonances = 0
    []
    [B10]
      awr = 9.927
      n_points = 20000
      seed = 10
      sigma_a0 = 3840
      sigma_p = 2.1
      resonances = 0
    []
    [Zr90]
      awr = 89.13
      n_points = 20000
      seed = 90
      sigma_a0 = 0.185
      sigma_p = 6.4
      resonances = 0
    []
    [He4]
      awr = 3.968
      n_points = 20000
      seed = 4
      sigma_a0 = 1e-6
      sigma_p = 0.76
      resonances = 0
    []
    [Fe56]
      awr = 55.45
      n_points = 20000
      seed = 56
      sigma_a0 = 2.56
      sigma_p = 11
      resonances = 0
    []
    [N14]
      awr = 13.88
      n_points = 20000
      seed = 14
      sigma_a0 = 1.9
      sigma_p = 10
      resonances = 0
    []
    [Si28]
      awr = 27.74
      n_points = 20000
      seed = 28
      sigma_a0 = 0.17
      sigma_p = 2.0
      resonances = 0
    []
  []
  [Materials]
    [air]
      id = 1
      nuclides = 'O16 N14'
      densities = '5.3067e-06 1.9991e-05'
    []
    [ss304]
      id = 2
      nuclides = 'Si28 Fe56'
      densities = '1.0322e-03 8.6799e-02'
    []
    [helium]
      id = 5
      nuclides = 'He4'
      densities = '2.4060e-04'
    []
    [zircaloy4]
      id = 7
      nuclides = 'O16 Fe56 Zr90'
      densities = '3.0816e-04 2.2393e-04 4.2862e-02'
    []
    [fuel16]
      id = 9
      nuclides = 'O16 U238 U235'
      densities = '4.6015e-02 2.2632e-02 3.7509e-04'
    []
    [fuel24]
      id = 10
      nuclides = 'O16 U238 U235'
      densities = '4.5948e-02 2.2416e-02 5.5824e-04'
    []
    [fuel31]
      id = 11
      nuclides = 'O16 U238 U235'
      densities = '4.5970e-02 2.2263e-02 7.2188e-04'
    []
    [borosilicate]
      id = 14
      nuclides = 'O16 Si28 B10'
      densities = '4.6609e-02 2.0080e-02 9.6474e-04'
    []
    [borated_water]
      id = 15
      nuclides = 'B10 H1 O16'
      densities = '7.9778e-06 4.9504e-02 2.4752e-02'
    []
  []
[]

[Executioner]
  type = Transient
[]

[Outputs]
  exodus = true
  csv = true
  perf_graph = false
[]

[UserObjects]
  inactive = 'tally univtally'
  [tally]
    type = OpenMCTally
    particle_type = 'neutron'
    estimator = 'COLLISION'
    scores = 'flux scatter (n,fission) 16'
    filters = 'energy particle'
    energy_bins = '1e-5 1e3 2e7'
    execute_on = 'initial'
  []
  [univtally]
    type = OpenMCTally
    particle_type = 'neutron'
    estimator = 'COLLISION'
    scores = 'kappa-fission'
    filters = 'universe'
    execute_on = 'initial'
  []
  [celltally]
    type = OpenMCTally
    id = 1
    particle_type = 'neutron'
    estimator = 'COLLISION'
    scores = 'kappa-fission'
    filters = 'cell'
    execute_on = 'initial'
  []
[]

[AuxVariables]
  [power]
    order = CONSTANT
    family = MONOMIAL
  []
[]

[AuxKernels]
  [cell_val]
    type = OpenMCTallyAux
    granularity = 'cell'
    score = 'kappa-fission'
    tally_id = 1
    execute_on = TIMESTEP_END
    variable = power
  []
[]
