C2H4-like polymer	! example material file: hydrogen lists no shell, its electrons are in the valence band
2		! how many elements are in this molecule of solid
6	2	! atomic number, contribution of the 1st element into the compound
1	4	! atomic number, contribution of the 2nd element into the compound
0.95	2000.0	5.0	7.0	! density [g/cm^3], speed of sound [m/s], Fermi level [eV], band gap [eV]
2		! number of shells of the first element: C
1	1	288.0d0	2	7.96d0	! number of CDF functions, shell designator, ionization potential, number of electrons, Auger time
300.0	157.65	100.0	! E0, A, Gamma coefficients
2	63	7.0d0	12	1.0d23	! the valence band: 8 electrons of carbon and 4 of hydrogen
15.0	135.11	8.0
22.0	202.67	12.0
0		! number of shells of the second element: H
