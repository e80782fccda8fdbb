Al2O3-like oxide	! example material file in the reference's full-CDF format
2		! how many elements are in this molecule of solid
13	2	! atomic number, contribution of the 1st element into the compound
8	3	! atomic number, contribution of the 2nd element into the compound
3.95	6510.0	4.4	8.8	! density [g/cm^3], speed of sound [m/s], Fermi level [eV], band gap [eV]
2		! number of shells of the first element: Al
2	2	80.0d0	8	27.27d0	! number of CDF functions, shell designator, ionization potential, number of electrons, Auger time
90.0	369.36	40.0	! E0, A, Gamma coefficients
130.0	221.62	60.0
4	63	8.8d0	24	1.0d23	! the valence band
14.0	115.55	3.0
24.0	269.62	8.0
33.0	269.62	12.0
55.0	115.55	20.0
1		! number of shells of the second element: O
2	1	538.0e0	2	4.613e0	! number of CDF functions, shell designator, ionization potential, number of electrons, Auger time
560.0	166.65	120.0
640.0	83.325	200.0
2		! phonon peaks:
0.09	0.002	0.006	! E0, A, Gamma coefficients
0.06	0.00005	0.002
