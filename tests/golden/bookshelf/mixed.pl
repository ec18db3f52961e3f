UCLA pl 1.0

a 0 0 : N
b 10 0 : N
macro 70 20 : N
pad 80 20 : N /FIXED
