"""Published embedded pairs the library does not ship, in the reference's tableau
struct format (a: stage rows, b, bt, p, pt), for the registration tests.

* Bogacki & Shampine, "A 3(2) pair of Runge-Kutta formulas", Appl. Math. Lett. 2
  (1989): 4 stages, first same as last.
* Cash & Karp, ACM TOMS 16 (1990): 6 stages, b of order 5, bt of order 4.
"""
BS23 = dict(
    a=[[1.0 / 2.0],
       [0.0, 3.0 / 4.0],
       [2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0]],
    b=[2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0, 0.0],
    bt=[7.0 / 24.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 8.0],
    p=3, pt=2)

CASH_KARP = dict(
    a=[[1.0 / 5.0],
       [3.0 / 40.0, 9.0 / 40.0],
       [3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0],
       [-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0],
       [1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0,
        253.0 / 4096.0]],
    b=[37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0],
    bt=[2825.0 / 27648.0, 0.0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0,
        1.0 / 4.0],
    p=5, pt=4)
