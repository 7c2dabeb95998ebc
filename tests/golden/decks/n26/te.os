rs =  0.1000000000000E+01

 computing  26 lattice sites  dimensionality 2 number density  0.31831E+00
  crystal type chosen by default to minimize vacancies 
   body centered cubic lattice
 q displs    0.00000   0.00000   0.50000
 q displs    0.50000
 npuc  2 nvacancies     6 box size 0.90378E+01 0.90378E+01
 number of cells in each direction    4    4
 nearest and next nearest neighbor distance   0.15977E+01 0.22594E+01

  shell 1mult  4 k   0.695213598
  shell 2mult  4 k   0.983180499
  shell 3mult  4 k   1.3904272
   spin  1 ke finite   13.5330145 infinite   13.
  fermiwv  3  1.3904272  1.41421356
  shell 1mult  4 k   0.695213598
  shell 2mult  4 k   0.983180499
  shell 3mult  4 k   1.3904272
   spin  2 ke finite   13.5330145 infinite   13.
  fermiwv  3  1.3904272  1.41421356
  free particle ke finite   1.04100112 infinite   0.999999999 difference 
  0.0410011171
  fermi wavevector =   1.41421356
  ideal square madelung constant -0.863102734 -0.863102734
  beginning fitpn n nk nf i0=  10 91 18 0
  ifcon ndim   1 2
  mnts lptable vol b0   501 501  81.681409  2.
  a pn np m mp i2 beta   4.51888839  1.57079633 3 7 6 2 -0.442586722  0.
  rms error   3.94776845E-07

 fitp m=   10    0 rcond   0.4904E-08
 poly=  -0.442587E+00 -0.130688E+00  0.935727E+00  0.689471E+00 -0.838311E+00
 poly=  -0.113999E+01  0.913706E+00
  k=     0.  0.695213598  0.983180499  1.3904272  1.55454486  1.966361
  2.08564079  2.19845843  2.50662827  2.78085439  2.8664391  2.9495415
  3.10908973  3.47606799  3.5449077  3.7438398  3.932722  4.05375705
  4.17128159
  f=    -0.158381763  0.0805122668  0.0301998041  0.00633103107  0.0029732184
  0.000179653949  1.91395395E-05 -3.85654324E-05 -2.66742448E-05
  2.51698735E-06  5.99169143E-06  7.27418001E-06  6.09464051E-06
  5.82740087E-08 -4.97191389E-07 -9.89558306E-07 -6.27977817E-07
 -3.28429961E-07 -1.10633176E-07
  computed madelung constant -0.86310248
